// The bonded launch: the CHARMM bonded terms and the PME exclusion correction of
// mdx_bonded_terms.h, for evaluations without a nonbonded launch to carry them.
#include "mdx_bonded_terms.h"

__global__ __launch_bounds__(BD_BLOCK) void k_bonded(BondedArgs a) { bonded_block(a, blockIdx.x); }

// arguments of the bonded terms of `terms` plus, with MDX_TERM_COUL_RECIP, the PME
// exclusion/self/background part; returns the number of workgroups they take
int mdx_bonded_args(const mdx_ctx* c, unsigned terms, BondedArgs* out, int block) {
  BondedArgs& a = *out;
  a.n = c->n;
  for (int q = 0; q < 3; ++q) {
    a.b.L[q] = c->box[q];
    a.b.inv[q] = 1.0 / c->box[q];
  }
  a.pos = c->pos;
  a.ffix = c->ffix.p;
  a.efix = c->efix.p;
  const bool on[4] = {(terms & MDX_TERM_BOND) != 0, (terms & MDX_TERM_ANGLE) != 0, (terms & MDX_TERM_DIHEDRAL) != 0,
                      (terms & MDX_TERM_IMPROPER) != 0};
  const int counts[4] = {c->nb, c->na, c->nd, c->ni};
  const int* idx[4] = {c->bond_idx.p, c->angle_idx.p, c->dih_idx.p, c->imp_idx.p};
  const double* par[4] = {c->bond_par.p, c->angle_par.p, c->dih_par.p, c->imp_par.p};
  int total = 0;
  for (int k = 0; k < 4; ++k) {
    a.cnt[k] = on[k] ? counts[k] : 0;
    a.blocks[k] = div_up(a.cnt[k], block);
    a.idx[k] = idx[k];
    a.par[k] = par[k];
    total += a.blocks[k];
  }
  a.refcompat = c->dih_refcompat;
  const bool recip = (terms & MDX_TERM_COUL_RECIP) != 0;
  a.cnt[4] = recip ? c->n_excl_pairs : 0;
  a.blocks[4] = recip ? std::max(1, div_up(c->n_excl_pairs, block)) : 0;
  a.idx[4] = c->excl_pairs.p;
  a.par[4] = nullptr;
  total += a.blocks[4];
  a.q = c->charge64.p;
  a.ke = 1.0 / (4.0 * M_PI * 0.5727652348434076);
  a.beta = c->beta;
  a.e_const = 0.0;
  if (recip) {
    const double V = c->box[0] * c->box[1] * c->box[2];
    const double e_self = -a.ke * c->beta / sqrt(M_PI) * c->q2sum;
    const double e_bg = -a.ke * M_PI * c->qsum * c->qsum / (2.0 * c->beta * c->beta * V);
    a.e_const = e_self + e_bg;
  }
  return total;
}

int mdx_bonded_enqueue(mdx_ctx* c, unsigned terms) {
  BondedArgs a;
  const int total = mdx_bonded_args(c, terms, &a);
  if (total == 0) return 0;
  k_bonded<<<total, BD_BLOCK, 0, c->stream>>>(a);
  MDX_LAUNCH_CHECK();
  return 0;
}
