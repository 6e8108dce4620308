"""ripley_hip: MI355X-native ripley Rectangle/Brick Q1 assembly + PASO
Jacobi-PCG, behind the escript operator interface.

    from ripley_hip import Brick, Poisson, SolverOptions
    dom = Brick(64, 64, 64)
    pde = Poisson(dom)
    pde.setValue(f=1.0, q=dom.whereZero(2))
    u = pde.getSolution()
"""
from ._lib import PasoException, RipleyHipError, LIB_PATH, load as load_library
from .core import Coefficient, Context, SolverOptions, SystemMatrix, get_context
from .domain import Brick, Rectangle
from .linearpde import IsotropicCoefficient, LameEquation, LinearPDE, Poisson, kronecker_scaled
from .gridio import (BYTEORDER_BIG_ENDIAN, BYTEORDER_LITTLE_ENDIAN, BYTEORDER_NATIVE, DATATYPE_FLOAT32,
                     DATATYPE_FLOAT64, DATATYPE_INT32, ContinuousFunction, Function, FunctionSpace,
                     ReducedFunction, Solution, readBinaryGrid, readBinaryGridFromZipped)

__all__ = ["Brick", "Rectangle", "LinearPDE", "Poisson", "LameEquation", "SolverOptions", "SystemMatrix",
           "Coefficient", "Context", "get_context", "kronecker_scaled", "IsotropicCoefficient",
           "PasoException", "RipleyHipError", "LIB_PATH", "load_library", "FunctionSpace", "Solution",
           "ContinuousFunction", "Function", "ReducedFunction", "readBinaryGrid", "readBinaryGridFromZipped",
           "BYTEORDER_NATIVE", "BYTEORDER_LITTLE_ENDIAN", "BYTEORDER_BIG_ENDIAN", "DATATYPE_INT32",
           "DATATYPE_FLOAT32", "DATATYPE_FLOAT64"]
