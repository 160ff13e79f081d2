"""Trajectory analysers (mdpy/analyser): RDF and coordination number with their pair-distance
histograms computed on the GPU, and the AnalyserResult container."""
from .analyser_result import AnalyserResult, load_analyser_result
from .rdf_analyser import RDFAnalyser
from .coordination_number_analyser import CoordinationNumberAnalyser

__all__ = ["AnalyserResult", "load_analyser_result", "RDFAnalyser", "CoordinationNumberAnalyser"]
