"""Helpers of mdpy/utils used by the analysers (topological atom selection)."""
from .select import (SELECTION_SUPPORTED_KEYWORDS, SELECTION_SUPPORTED_TOPOLOGICAL_KEYWORDS,
                     check_selection_condition, check_topological_selection_condition, parse_selection_condition,
                     select)

__all__ = ["select", "check_selection_condition", "check_topological_selection_condition",
           "parse_selection_condition", "SELECTION_SUPPORTED_KEYWORDS", "SELECTION_SUPPORTED_TOPOLOGICAL_KEYWORDS"]
