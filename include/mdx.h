/*
 * mdx.h — C-ABI of libmdx.so, the MI355X (gfx950) force-and-integrate step behind mdpy's
 * Constraint / Integrator / Simulation API.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative MDX_ERR_* code; mdx_last_error()
 *     returns a thread-local message for the last failure.
 *   - Units are mdpy's internal units: A, Da, fs, e, K; energy Da*A^2/fs^2
 *     (mdpy/unit/__init__.py:31-41).
 *   - "d_" pointers are device pointers (HIP global memory) owned by the caller (PyTorch
 *     tensors); the library only borrows them. "h_" pointers are host memory read during
 *     the call. The library owns its scratch (cell tables, neighbour lists, PME grids,
 *     fixed-point accumulators) inside the context.
 *   - All device work is enqueued on the context's stream; nothing synchronises except
 *     functions whose name ends in _fetch or that return host-visible results (documented).
 *   - One context per host thread; contexts are independent.
 *   - Positions and velocities are float64 N x 3 row-major (x0,y0,z0,x1,...).
 *   - Orthorhombic boxes only (edge lengths); the reference's CUDA kernels make the same
 *     restriction (mdpy/constraint/charmm_nonbonded_constraint.py:145-150).
 */
#ifndef MDX_H
#define MDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDX_OK                 0
#define MDX_ERR_INVALID       -1   /* bad argument or call order */
#define MDX_ERR_HIP           -2   /* a HIP runtime call failed */
#define MDX_ERR_OVERFLOW      -3   /* cell or neighbour-list capacity exceeded */
#define MDX_ERR_PARTICLE_LOSS -4   /* an atom moved >= 2 box images (mdpy ParticleLossError) */
#define MDX_ERR_NONFINITE     -5   /* NaN/Inf positions or forces detected on the device */
#define MDX_ERR_CELL_CUTOFF   -6   /* cutoff too large for the box (mdpy CellListPoorDefinedError) */

/* force terms (bit mask) */
#define MDX_TERM_LJ            (1u << 0)  /* CHARMM 12-6 LJ, r <= rc (CharmmNonbondedConstraint) */
#define MDX_TERM_COUL_DIRECT   (1u << 1)  /* PME direct space: q_i q_j erfc(beta r)/r, r <= rc   */
#define MDX_TERM_COUL_RECIP    (1u << 2)  /* PME reciprocal + self + exclusion + background       */
#define MDX_TERM_BOND          (1u << 3)
#define MDX_TERM_ANGLE         (1u << 4)
#define MDX_TERM_DIHEDRAL      (1u << 5)
#define MDX_TERM_IMPROPER      (1u << 6)
#define MDX_TERM_COUL_PLAIN    (1u << 7)  /* plain Coulomb, all pairs, minimum image, no cutoff  */
#define MDX_TERM_MLP           (1u << 8)  /* learned surrogate of LJ + PME (mdx_mlp_*)          */
#define MDX_NTERMS             9
/* energy slot of the kinetic energy in mdx_energies_fetch output (after the MDX_NTERMS terms) */
#define MDX_E_KINETIC          MDX_NTERMS

/* neighbour-set modes */
#define MDX_PAIRS_EXACT        0  /* every minimum-image pair with r <= rc */
#define MDX_PAIRS_REFERENCE    1  /* plus mdpy's cell-stencil restriction (cell_list.py:74-116) */

/* integrators */
#define MDX_INT_VELOCITY_VERLET 0 /* textbook velocity Verlet (NVE)                           */
#define MDX_INT_LANGEVIN        1 /* G-JF Langevin with mdpy's a, b, sigma                    */
#define MDX_INT_VERLET_REF      2 /* verlet_integrator.py:20-50 as written                    */
#define MDX_INT_LANGEVIN_REF    3 /* langevin_integrator.py:37-72 as written                  */

typedef struct mdx_ctx mdx_ctx;
typedef struct mdx_mlp mdx_mlp;

/* ---- library / context ------------------------------------------------------------------ */
const char* mdx_last_error(void);
/* 1/(4 pi eps0) in internal units, the Coulomb factor every kernel uses (electrostatic_constraint.py:21,60). */
double mdx_coulomb_constant(void);
int mdx_version(void);
/* Create a context on HIP device `device`, enqueuing on `hip_stream` (NULL = default stream). */
int mdx_ctx_create(int device, void* hip_stream, mdx_ctx** out);
int mdx_ctx_destroy(mdx_ctx* ctx);
int mdx_ctx_set_stream(mdx_ctx* ctx, void* hip_stream);

/* ---- system (State / Topology.join, mdpy/core/state.py:19-28, topology.py:59-80) --------- */
/* Bind caller-owned device arrays: positions (N*3 f64), velocities (N*3 f64). `h_box` = 3 edges. */
int mdx_system_set(mdx_ctx* ctx, int n_atoms, double* d_pos, double* d_vel, const double* h_box);
int mdx_set_box(mdx_ctx* ctx, const double* h_box);
/* Per-atom data from host arrays (copied): masses (N), charges (N, e), LJ (N*4: eps, sigma,
 * eps14, sigma14 — CharmmNonbondedConstraint.bind_ensemble, charmm_nonbonded_constraint.py:48-62). */
int mdx_set_masses(mdx_ctx* ctx, const double* h_mass);
int mdx_set_charges(mdx_ctx* ctx, const double* h_charge);
int mdx_set_lj(mdx_ctx* ctx, const double* h_lj);
/* Exclusions (1-2 + 1-3, topology.py:122-167) and 1-4 partners (topology.py:188-200) as CSR
 * (ptr has N+1 entries). */
int mdx_set_exclusions(mdx_ctx* ctx, const int* h_ptr, const int* h_idx);
int mdx_set_scaling(mdx_ctx* ctx, const int* h_ptr, const int* h_idx);
/* Bonded terms (bind_ensemble of charmm_{bond,angle,dihedral,improper}_constraint.py):
 * bonds nb x 2 ids + nb x 2 (k, r0); angles na x 3 + na x 4 (k, theta0, k_ub, u0);
 * dihedral terms nd x 4 + nd x 3 (k, n, delta); impropers ni x 4 + ni x 2 (k, psi0).
 * dihedral_refcompat != 0 reproduces force_val = -k(1 - n sin(n phi - delta))
 * (charmm_dihedral_constraint.py:81) instead of -dE/dphi. */
int mdx_set_bonds(mdx_ctx* ctx, int nb, const int* h_idx, const double* h_par);
int mdx_set_angles(mdx_ctx* ctx, int na, const int* h_idx, const double* h_par);
int mdx_set_dihedrals(mdx_ctx* ctx, int nd, const int* h_idx, const double* h_par, int dihedral_refcompat);
int mdx_set_impropers(mdx_ctx* ctx, int ni, const int* h_idx, const double* h_par);

/* ---- State.set_positions: wrap + reference cell list (pbc.py:28-36, cell_list.py:74-116) ---- */
/* Wrap d_pos in place with x + (-rint(x * inv_L)) * L; returns MDX_ERR_PARTICLE_LOSS (after a
 * sync) if any |shift| >= 2. Synchronises. */
int mdx_wrap_positions(mdx_ctx* ctx);
/* Reference binning of the current positions with the cell-list cutoff `cell_cutoff`:
 * n = max(3, floor((L + 2)/rc)), w = (L + 2)/n, idx = floor((x - min x)/w) (bit-exact float64).
 * Writes d_cell_index (N*3 int32) and returns grid dims and the largest cell population in
 * h_n[3], h_maxpc. Synchronises. */
int mdx_celllist_reference(mdx_ctx* ctx, double cell_cutoff, int* d_cell_index, int* h_n, int* h_maxpc);
/* The caller wrote the bound position array directly (not through this library): the next
 * mdx_integrate evaluates forces at the new positions before its first step. Without it, G-JF and
 * velocity-Verlet calls start from the forces and energies of the previous call's last step (the
 * integrator's cache, as integrator.py:17-22 keeps positions between calls); every other entry
 * point of this library that changes positions, parameters or the force buffers drops that cache
 * by itself. No counterpart in the reference, whose integrate() always re-evaluates. */
int mdx_positions_changed(mdx_ctx* ctx);
/* mdx_celllist_reference without a context and without any device allocation. No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code around stream
 * capture, tight serving loops). Positions d_pos (n x 3 f64, device) in a box of edges h_box;
 * enqueued on `hip_stream` (NULL = default stream); writes d_cell_index (n x 3 int32) and returns
 * the grid in h_n[3] and the largest cell population in h_maxpc (synchronises). Its reductions
 * and per-cell counts use one fixed device array per GPU (held until the process ends) of up to
 * 16,384 cells: larger grids return MDX_ERR_INVALID. Calls that overlap on one GPU race on that
 * array; each call clears what it uses first. */
int mdx_celllist_reference_noalloc(void* hip_stream, int n, const double* d_pos, const double* h_box,
                                   double cell_cutoff, int* d_cell_index, int* h_n, int* h_maxpc);
/* Dense cell table [nx][ny][nz][maxpc], atoms ascending within a cell, -1 padded (cell_list.py:105-116),
 * for the binning of the last mdx_celllist_reference call. */
int mdx_celllist_table(mdx_ctx* ctx, int* d_table, int maxpc);

/* ---- nonbonded (CharmmNonbondedConstraint, charmm_nonbonded_constraint.py:64-226) -------- */
/* cutoff rc, list skin (0 = rebuild every evaluation), pair set (MDX_PAIRS_*), cell_cutoff for
 * the reference stencil (the reference cell list's cutoff, 12 A by default). */
int mdx_nonbonded_config(mdx_ctx* ctx, double cutoff, double skin, int pair_mode, double cell_cutoff);
/* Enumerate the unique neighbour pairs (i<j, original ids) the configured pair set contains at
 * the current positions, decided in float64. d_pairs holds capacity x 2 int32. Returns the pair
 * count in *h_count (may exceed capacity: then only capacity pairs are written). Synchronises. */
int mdx_neighbor_pairs(mdx_ctx* ctx, int* d_pairs, int64_t capacity, int64_t* h_count);

/* ---- PME (new; no reference counterpart, mdpy/constraint/__init__.py:22) ------------------ */
/* grid K (3 ints), B-spline order (4), Ewald beta (1/A). */
int mdx_pme_config(mdx_ctx* ctx, const int* h_grid, int order, double beta);
/* Reciprocal space in one launch with two grid barriers (enable != 0) or in three launches
 * (spread + 2D FFT, x-axis FFT, 2D inverse + gather; the default). Same forces. The one-launch
 * kernel needs its Kx blocks resident together and is never used inside an mdx_batch; on
 * MI355X its barriers (cache write-back and invalidation across the XCDs) cost more than the
 * launch gaps they remove, so it stays off unless selected. */
int mdx_pme_set_fused(mdx_ctx* ctx, int enable);
/* Phase launches (enable != 0; no reference counterpart): force evaluations of LJ + PME run the
 * three reciprocal launches on the context's stream, each also running a share of the
 * nonbonded workgroups (frac0 of them with the spread/2D FFT launch, frac1 with the x-axis
 * launch, the rest with the gather launch, which also carries the bonded terms) instead of the
 * nonbonded kernel on the main stream and the reciprocal chain on a side stream joined after.
 * Same forces as the two-stream path up to float32 summation order within an atom's partial
 * sums (the nonbonded workgroups are wider). Ignored where the plane PME path is not used. */
int mdx_force_phases_config(mdx_ctx* ctx, int enable, double frac0, double frac1);

/* ---- force evaluation (Constraint.update / Ensemble.update, ensemble.py:53-61) ------------ */
/* Evaluate the terms in `terms` at the current positions. If d_force_out != NULL, write the
 * summed forces (N*3 f64). Energies accumulate on the device; read them with mdx_energies_fetch. */
int mdx_compute_forces(mdx_ctx* ctx, unsigned terms, double* d_force_out);
/* As mdx_compute_forces, with the kernels that integrator steps recording no energies run (the
 * same forces to float32 rounding; the direct-space Coulomb force uses a tabulated Ewald
 * correction). The energy accumulators are left at zero. */
int mdx_compute_forces_only(mdx_ctx* ctx, unsigned terms, double* d_force_out);
/* Copy the per-term energies of the last evaluation (+ kinetic energy at MDX_E_KINETIC) to
 * h_out[MDX_NTERMS + 1]. Synchronises. */
int mdx_energies_fetch(mdx_ctx* ctx, double* h_out);

/* ---- integration (Integrator.integrate, the modules of mdpy/integrator) ------------------ */
int mdx_integrator_config(mdx_ctx* ctx, int kind, double dt, double gamma, double kT, uint64_t seed);
/* Forget integrator history (Integrator.erase_cache, integrator.py:17-22). */
int mdx_integrator_reset(mdx_ctx* ctx);
/* Use caller-provided standard normals d_noise[nsteps][N][3] (f64) instead of the built-in
 * Philox4x32-10 stream, step by step from the next step on (NULL switches back). */
int mdx_set_external_noise(mdx_ctx* ctx, const double* d_noise, int nsteps);
/* Advance `nsteps` steps with force terms `terms`. use_graph != 0 replays the steps after the
 * first from hipGraphs of fixed sizes (forces-only runs of 32/16/4/1 steps and a single energy
 * step), captured once per (terms, integrator kind) and reused by calls of any length.
 * Positions/velocities in the bound arrays are updated (positions wrapped as State.set_positions
 * does); per-step energies go to the energy history. The call's last step always evaluates
 * energies. */
int mdx_integrate(mdx_ctx* ctx, unsigned terms, int nsteps, int use_graph);
/* Evaluate potential energies every `every` steps during mdx_integrate (default 1; forces are
 * evaluated every step). The energy history then holds rows for steps 0, every, 2*every, ... */
int mdx_set_energy_interval(mdx_ctx* ctx, int every);
/* Energies of the last mdx_integrate call, one row per energy step (steps 0, every, ...):
 * h_out[k*(MDX_NTERMS+1) + t] for k < min(rows, h_cap); *h_count = rows. Synchronises. */
int mdx_energy_history_fetch(mdx_ctx* ctx, double* h_out, int h_cap, int* h_count);
/* As mdx_energy_history_fetch without synchronising: the copy of min(rows, h_cap) rows into h_out
 * (pinned host memory for an asynchronous copy) is enqueued on the context's stream, ordered
 * after the integrate call; *h_count = the rows copied. The rows are valid in h_out once the
 * caller's stream has reached this point (an event recorded after the call). */
int mdx_energy_history_copy_async(mdx_ctx* ctx, double* h_out, int h_cap, int* h_count);
/* Status flags raised on the device since the last call (bit 0 overflow, 1 particle loss,
 * 2 non-finite, 3 PME coverage: an atom missed by the reciprocal-space gather, 4 the fused
 * PME kernel's grid barrier gave up waiting for its blocks); clears them. Synchronises. */
int mdx_status_fetch(mdx_ctx* ctx, unsigned* h_flags);

/* ---- energy minimisation (mdpy/minimizer) ----------------------------------------------- */
/* SteepestDescentMinimizer.minimize (steepest_descent_minimizer.py:30-53): x_i <- x_i +
 * alpha F_i/|F_i| (wrapped) until |(E - E_prev)/E_prev| < energy_tol or max_iter iterations; every
 * evaluation is the fused evaluation of `terms` (energies included). Returns the iterations done
 * and the final potential energy. Synchronises once per iteration (energy read-back). An atom
 * with no force does not move (the reference divides by its zero norm). */
int mdx_minimize_sd(mdx_ctx* ctx, unsigned terms, double alpha, double energy_tol, int max_iter, int* h_iters,
                    double* h_energy);
/* ConjugateGradientMinimizer.minimize (conjugate_gradient_minimizer.py:33-77) as written: up to
 * max_sub_iter finite-difference probes of step theta along the current forces per iteration
 * (each from the previous probe's positions), the last probe's step alpha t applied to the
 * iteration's starting positions; the probe loop ends early when |F + alpha omega|^2/|F0|^2 <
 * force_tol. Sums run on the device in a fixed order (deterministic). */
int mdx_minimize_cg(mdx_ctx* ctx, unsigned terms, double theta, double force_tol, int max_sub_iter, double energy_tol,
                    int max_iter, int* h_iters, double* h_energy);

/* ---- analysis (mdpy/analyser/rdf_analyser.py:34-85, coordination_number_analyser.py:34-80) ---- */
/* For each atom i = d_sel1[a] and every j in d_sel2, the minimum-image distance in float64 of
 * frames d_pos[n_frames][n_atoms][3] in a box of edges h_box, binned on the nbins + 1 edges
 * d_edges (numpy.linspace(0, rc, nbins + 1)) exactly as numpy.histogram bins it; d_counts
 * (n1 x nbins int64, zeroed by the caller) accumulates over frames. Enqueued on hip_stream. */
int mdx_pair_histogram(void* hip_stream, int n_frames, int n_atoms, const double* d_pos, const double* h_box,
                       const int* d_sel1, int n1, const int* d_sel2, int n2, const double* d_edges, int nbins,
                       int64_t* d_counts);

/* ---- batched replicas (new; no reference counterpart) ---------------------------------- */
/* Many independent systems, one context each (same device), advanced together: each context
 * keeps its own box, lists, PME grid, integrator and graph. Created on the caller's stream
 * `hip_stream` (NULL = default); the contexts are borrowed and must outlive the batch. */
typedef struct mdx_batch mdx_batch;
#define MDX_BATCH_ONE_GRAPH   0  /* one hipGraph with a branch per replica (disabled: returns
                                    MDX_ERR_INVALID until a host crash is understood)          */
#define MDX_BATCH_STREAMS     1  /* each replica replays its own graph on its own stream         */
#define MDX_BATCH_SEQUENTIAL  2  /* replica after replica (the per-replica loop, for comparison) */
#define MDX_BATCH_FUSED       3  /* one launch per kernel per step covering every replica's
                                    workgroups (prologue, sort/fill, PME planes, list build,
                                    nonbonded + bonded), the steps captured in hipGraphs. Needs
                                    G-JF or velocity Verlet, LJ + PME (+ bonded terms), the PME
                                    plane path, a list skin > 0 and the same lanes per atom in
                                    every replica; trajectories equal single-context runs bit
                                    for bit. No replica-count limit beyond device memory.       */
int mdx_batch_create(mdx_ctx** ctxs, int n_ctx, void* hip_stream, mdx_batch** out);
/* Advance every replica `nsteps` steps of `terms` (each with its own integrator settings; all
 * replicas need the same energy interval). Energy histories are per context, as after
 * mdx_integrate. */
int mdx_batch_integrate(mdx_batch* batch, unsigned terms, int nsteps, int mode);
/* Average ms per launch of one batched kernel of MDX_BATCH_FUSED, alone over every replica (HIP
 * events on the batch's stream), after a fused integrate call: which 0 nonbonded (forces only),
 * 4 step prologue (the replicas' state saved and restored), 5 list build of a rebuild step,
 * 6 / 7 / 8 PME plane forward / x-axis pass / inverse + gather. The next call re-evaluates the
 * replicas' forces. */
int mdx_batch_time_kernel(mdx_batch* batch, int which, int reps, double* h_avg_ms);
int mdx_batch_destroy(mdx_batch* batch);

/* ---- learned surrogate (new; no reference counterpart) ---------------------------------- */
/* A per-atom MLP (descriptors within rd of each atom: 32 radial Gaussians + 16 angular moment
 * invariants per neighbour element; per element 96 -> 256 -> 256 -> 256 -> 1, SiLU, BF16 on
 * MFMA) whose energy gradient stands in for LJ + PME. h_elem: element per atom (0 = O, 1 = H).
 * The surrogate keeps its own rd (+ skin) neighbour list on the host context's positions and
 * runs on the host context's stream. Parameters per element: mdx_mlp_param_count() floats
 * (W1 96x256 | b1 | W2 256x256 | b2 | W3 256x256 | b3 | w4 256 | b4), row-major. */
int mdx_mlp_create(mdx_ctx* ctx, const int* h_elem, double rd, double skin, mdx_mlp** out);
int mdx_mlp_destroy(mdx_mlp* mlp);
int mdx_mlp_param_count(void);
int mdx_mlp_set_params(mdx_mlp* mlp, int elem, const float* h_params);
int mdx_mlp_get_params(mdx_mlp* mlp, int elem, float* h_params);
/* feature normalisation x = (G - mean) / std (96 each) */
int mdx_mlp_set_norm(mdx_mlp* mlp, const float* h_mean, const float* h_std);
/* descriptors G (N x 96 f32, device, original atom order) at the host context's positions */
int mdx_mlp_descriptors(mdx_mlp* mlp, float* d_out);
/* surrogate forces (N x 3 f64, device; may be NULL) and total energy (host; may be NULL) */
int mdx_mlp_eval(mdx_mlp* mlp, double* d_force_out, double* h_energy);
/* attach to the host context: MDX_TERM_MLP in mdx_compute_forces / mdx_integrate then adds the
 * surrogate's forces (energy slot MDX_TERM_MLP's bit index). NULL detaches. */
int mdx_ctx_set_mlp(mdx_ctx* ctx, mdx_mlp* mlp);
/* One force-matching step (loss = mean squared force error per component) on one frame
 * (d_pos, d_fref: N x 3 f64 device). lr > 0 applies an Adam update; the loss goes to *h_loss
 * when not NULL (synchronises). */
int mdx_mlp_train_step(mdx_mlp* mlp, const double* d_pos, const double* d_fref, double lr, double* h_loss);
int mdx_mlp_adam_config(mdx_mlp* mlp, double beta1, double beta2, double eps);
/* gradients of the last mdx_mlp_train_step (same layout as the parameters) */
int mdx_mlp_get_grads(mdx_mlp* mlp, int elem, float* h_grad);
/* C (M x N f32) = A (M x K bf16) B (K x N bf16), all row-major device arrays: the MFMA GEMM
 * the dense layers use, exposed for testing */
int mdx_mlp_gemm_check(mdx_mlp* mlp, int M, int N, int K, const void* d_A, const void* d_B, float* d_C);

/* ---- instrumentation ------------------------------------------------------------------- */
/* Number of neighbour-list rebuilds and evaluations since the context was created. */
int mdx_stats_fetch(mdx_ctx* ctx, int64_t* h_rebuilds, int64_t* h_evals);
/* Average duration (ms) of one launch of a hot kernel, timed with HIP events on the context's
 * stream: which 0 = nonbonded kernel (current list, forces only), 1 = force evaluation of
 * `terms`, 2 = PME reciprocal chain, 3 = one integrator step of `terms`, 4 = the fused step
 * prologue alone (state saved and restored), 5 = the neighbour-list build of a rebuild step (the
 * next evaluation rebuilds in full), 6 / 7 / 8 = the PME plane forward kernel / x-axis pass /
 * inverse + gather kernel alone. Synchronises. */
int mdx_time_kernel(mdx_ctx* ctx, int which, unsigned terms, int reps, double* h_avg_ms);
/* Lanes per atom of the nonbonded kernel: 16, 32 or 64 set directly, or 0 = time each variant
 * on the current neighbour list (`reps` forces-only launches of `terms`' LJ/direct parts) and
 * keep the fastest. The choice goes to *h_lanes (may be NULL). Synchronises. */
int mdx_nonbonded_tune(mdx_ctx* ctx, unsigned terms, int lanes, int reps, int* h_lanes);
/* Resident nonbonded workgroups (no reference counterpart): per_cu > 0 launches per_cu
 * workgroups per compute unit that loop over the nonbonded work (default 3, the measured best
 * beside the concurrent PME kernels); 0 launches one workgroup per unit of work. Same forces
 * either way. */
int mdx_nonbonded_persist(mdx_ctx* ctx, int per_cu);
/* Scheduled list rebuilds (no reference counterpart): every > 0 also rebuilds the neighbour
 * list in the integrator step whose step counter is a multiple of `every`, besides the rebuilds
 * the displacement test asks for (which still happen whenever an atom moves more than skin/2);
 * 0 (default) rebuilds on the displacement test only. Replicas of a batch that share the
 * interval rebuild together, so a batched step holds either all their list builds or none.
 * Applies to the fused step prologue (G-JF and velocity Verlet on a list grid). */
int mdx_nonbonded_rebuild_every(mdx_ctx* ctx, int every);
/* Cell size of the neighbour-list grid (no reference counterpart; the reference's own cell list,
 * mdpy/core/cell_list.py, keeps its cutoff-wide cells): per_radius = 1, 2 (default) or 3 cells
 * per list radius r_c + skin along each box edge, so cells at least (r_c + skin)/per_radius
 * wide. Smaller cells scan fewer candidates outside the list radius but visit more stencil
 * cells. The lists hold the same pairs for every choice, rows in another order: forces and
 * energies agree to float32 summation rounding (each row's float32 partial sums follow its
 * order). The next evaluation bins the atoms again and rebuilds. */
int mdx_nonbonded_cells(mdx_ctx* ctx, int per_radius);
/* Candidate (list) pair slots and in-cutoff unique pairs of the last nonbonded evaluation
 * (counts the list on the device; for roofline accounting; synchronises). */
int mdx_pair_counts(mdx_ctx* ctx, int64_t* h_list_slots, int64_t* h_unique_in_cutoff);

#ifdef __cplusplus
}
#endif
#endif /* MDX_H */
