"""Per-atom pair-distance histograms of a trajectory on the GPU (libmdx mdx_pair_histogram)."""
import ctypes

import numpy as np

from .. import _lib as L


def pair_histograms(trajectory, ids1, ids2, cutoff_radius, num_bins, device=0):
    """int64 counts (len(ids1), num_bins) summed over frames: for each atom i of ids1 the
    minimum-image distances to the atoms of ids2 binned on numpy.linspace(0, rc, num_bins + 1)
    as numpy.histogram bins them; and the bin edges."""
    import torch
    if not torch.cuda.is_available():
        from ..error import DeviceError
        raise DeviceError("the analysers' pair histograms run on the GPU")
    dev = torch.device("cuda", device)
    edges = np.linspace(0, cutoff_radius, num_bins + 1)
    pos = torch.as_tensor(np.ascontiguousarray(trajectory.positions, dtype=np.float64), device=dev)
    s1 = torch.as_tensor(np.asarray(ids1, dtype=np.int32), device=dev)
    s2 = torch.as_tensor(np.asarray(ids2, dtype=np.int32), device=dev)
    e = torch.as_tensor(edges, device=dev)
    counts = torch.zeros((len(ids1), num_bins), dtype=torch.int64, device=dev)
    box = np.diag(np.asarray(trajectory.pbc_matrix, dtype=np.float64)).copy()
    stream = torch.cuda.current_stream(dev).cuda_stream
    for f0 in range(0, trajectory.num_frames, 65535):
        nf = min(65535, trajectory.num_frames - f0)
        L.call("mdx_pair_histogram", ctypes.c_void_p(stream), nf, trajectory.topology.num_particles,
               ctypes.c_void_p(pos[f0].data_ptr()), (ctypes.c_double * 3)(*box), ctypes.c_void_p(s1.data_ptr()),
               len(ids1), ctypes.c_void_p(s2.data_ptr()), len(ids2), ctypes.c_void_p(e.data_ptr()), num_bins,
               ctypes.c_void_p(counts.data_ptr()))
    return counts.cpu().numpy(), edges
