"""MI355X-native (gfx950) 2D RCWA unit-cell solver.

Drop-in for the reference toolbox's ``orf.rcwa2D`` path: the structure
builders mirror toolbox/+orf's model functions, and ``rcwa2D``, ``rcwa1D``,
``buildPhaseLibrary`` and ``convergeN`` run every solve through the HIP
library liborf_rcwa.so (C ABI in include/orf_rcwa.h).
"""
from .structure import (OrfError, OrfWarning, alignedSampleCount, cellGrid, layerEps, makeLayer, makeStructure, material,
                        materialConst, shapeCircle, shapeCross, shapePolygon, shapeRect, slantedLayers)
from .rcwa import Plan, Problem, buildPhaseLibrary, convergeN, rcwa1D, rcwa2D, rcwaFields

__all__ = ["OrfError", "OrfWarning", "alignedSampleCount", "cellGrid", "layerEps", "makeLayer", "makeStructure", "material",
           "materialConst", "shapeCircle", "shapeCross", "shapePolygon", "shapeRect", "slantedLayers",
           "Plan", "Problem", "buildPhaseLibrary", "convergeN", "rcwa1D", "rcwa2D", "rcwaFields"]
