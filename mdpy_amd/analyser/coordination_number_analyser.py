"""Coordination number between two selections (mdpy/analyser/coordination_number_analyser.py:
17-105): the cumulative pair-distance histogram per atom of the first selection, averaged over
frames; mean and std over the first selection. Histograms from the GPU (pair_histograms)."""
import numpy as np

from ..utils import check_topological_selection_condition, parse_selection_condition, select
from ..unit import check_quantity_value
from .analyser_result import AnalyserResult
from .pair_histogram import pair_histograms


class CoordinationNumberAnalyser:
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
        # the frame mean of the per-frame cumulative sums is the cumulative sum of the summed counts
        # over the frame count
        cn = np.cumsum(counts, axis=1) / trajectory.num_frames
        mean_cn, std_cn = cn.mean(0), cn.std(0)
        mean_cn[0], std_cn[0] = mean_cn[1], std_cn[1]  # the self distance lands in bin 0
        title = "Coordination number function between %s --- %s" % (
            parse_selection_condition(self._selection_condition_1),
            parse_selection_condition(self._selection_condition_2))
        description = {
            "mean": "The mean value of coordination number function, unit: dimensionless",
            "std": "The std value of coordination number function, unit: dimensionless",
            "cutoff_radius": "The cutoff radius of coordination number function, unit: default_length_unit",
            "num_bins": "The number of bins used to construct coordination number curve, unit: dimensionless",
            "bin_edge": "The bin edge of coordination number function, unit: default_length_unit",
        }
        data = {"mean": mean_cn, "std": std_cn, "cutoff_radius": self._cutoff_radius, "num_bins": self._num_bins,
                "bin_edge": bin_edge}
        return AnalyserResult(title=title, description=description, data=data)

    selection_condition_1 = property(lambda self: self._selection_condition_1)
    selection_condition_2 = property(lambda self: self._selection_condition_2)
    cutoff_radius = property(lambda self: self._cutoff_radius)
    num_bins = property(lambda self: self._num_bins)
