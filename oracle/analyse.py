"""Float64 restatement of mdpy's RDF and coordination-number analysis loops -- TEST
INFRASTRUCTURE ONLY (see oracle/__init__.py). positions: (frames, N, 3); box: 3 edges."""
import numpy as np

from .mdref import unwrap_diag


def _dist(positions, frame, i, ids2, box):
    # rdf_analyser.py:48-53: vec = x_i - x_j, minimum image, r = sqrt(sum vec^2)
    vec = unwrap_diag(positions[frame, i, :] - positions[frame, ids2, :], box)
    return np.sqrt((vec ** 2).sum(1))


def pair_counts(positions, box, ids1, ids2, rc, nbins):
    """Per atom of ids1: numpy.histogram counts over [0, rc], summed over frames (int64)."""
    out = np.zeros((len(ids1), nbins), dtype=np.int64)
    edges = None
    for a, i in enumerate(ids1):
        for f in range(positions.shape[0]):
            h, edges = np.histogram(_dist(positions, f, i, ids2, box), nbins, [0, rc])
            out[a] += h
    return out, edges


def rdf(positions, box, ids1, ids2, rc, nbins):
    """rdf_analyser.py:41-65: (mean, std, bin_edge)."""
    hist, edge = [], None
    for i in ids1:
        per = []
        for f in range(positions.shape[0]):
            h, edge = np.histogram(_dist(positions, f, i, ids2, box), nbins, [0, rc])
            per.append(h)
        hist.append(np.vstack(per).mean(0))
    width = edge[1] - edge[0]
    hist = np.vstack(hist) / (4 * np.pi * edge[1:] ** 2 * width)
    mean, std = hist.mean(0), hist.std(0)
    factor = mean.mean()
    mean /= factor
    std /= factor
    mean[0], std[0] = mean[1], std[1]
    return mean, std, edge


def coordination_number(positions, box, ids1, ids2, rc, nbins):
    """coordination_number_analyser.py:41-60: (mean, std, bin_edge)."""
    cn, edge = [], None
    for i in ids1:
        per = []
        for f in range(positions.shape[0]):
            h, edge = np.histogram(_dist(positions, f, i, ids2, box), nbins, [0, rc])
            per.append(np.cumsum(h))
        cn.append(np.vstack(per).mean(0))
    cn = np.vstack(cn)
    mean, std = cn.mean(0), cn.std(0)
    mean[0], std[0] = mean[1], std[1]
    return mean, std, edge
