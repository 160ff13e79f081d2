"""Radial distribution function between two selections (mdpy/analyser/rdf_analyser.py:17-110).

The per-atom, per-frame distance histograms come from the GPU (pair_histograms); averaging
over frames, the shell-volume normalisation, the mean / std over the first selection and the
overall scaling are the reference's (rdf_analyser.py:44-65)."""
import numpy as np

from ..utils import check_topological_selection_condition, parse_selection_condition, select
from ..unit import check_quantity_value
from .analyser_result import AnalyserResult
from .pair_histogram import pair_histograms


class RDFAnalyser:
    def __init__(self, selection_condition_1, selection_condition_2, cutoff_radius, num_bins: int) -> None:
        check_topological_selection_condition(selection_condition_1)
        check_topological_selection_condition(selection_condition_2)
        self._selection_condition_1 = selection_condition_1
        self._selection_condition_2 = selection_condition_2
        self._cutoff_radius = float(check_quantity_value(cutoff_radius))
        if not isinstance(num_bins, int):
            raise TypeError("num_bins must be an int, got %s" % type(num_bins))
        self._num_bins = num_bins

    def analysis(self, trajectory, is_dimensionless=True) -> AnalyserResult:
        ids1 = select(trajectory, self._selection_condition_1)[0]
        ids2 = select(trajectory, self._selection_condition_2)[0]
        counts, bin_edge = pair_histograms(trajectory, ids1, ids2, self._cutoff_radius, self._num_bins)
        hist = counts / trajectory.num_frames  # frame mean per atom of selection 1
        bin_width = bin_edge[1] - bin_edge[0]
        hist = hist / (4 * np.pi * bin_edge[1:] ** 2 * bin_width)
        mean_hist, std_hist = hist.mean(0), hist.std(0)
        factor = mean_hist.mean()
        mean_hist /= factor
        std_hist /= factor
        mean_hist[0], std_hist[0] = mean_hist[1], std_hist[1]  # the self distance lands in bin 0
        title = "RDF between %s --- %s" % (parse_selection_condition(self._selection_condition_1),
                                           parse_selection_condition(self._selection_condition_2))
        description = {
            "mean": "The mean value of RDF function, unit: dimensionless",
            "std": "The std value of RDF function, unit: dimensionless",
            "cutoff_radius": "The cutoff radius of RDF function, unit: default_length_unit",
            "num_bins": "The number of bins used to construct RDF curve, unit: dimensionless",
            "bin_edge": "The bin edge of RDF function, unit: default_length_unit",
        }
        data = {"mean": mean_hist, "std": std_hist, "cutoff_radius": self._cutoff_radius,
                "num_bins": self._num_bins, "bin_edge": bin_edge}
        return AnalyserResult(title=title, description=description, data=data)

    selection_condition_1 = property(lambda self: self._selection_condition_1)
    selection_condition_2 = property(lambda self: self._selection_condition_2)
    cutoff_radius = property(lambda self: self._cutoff_radius)
    num_bins = property(lambda self: self._num_bins)
