"""MI355X numeric phase of the Sparspak.jl sparse direct solver.

Public surface (mirrors the reference's API, Julia bangs dropped):
    Problem, SparseSolver, SparseSpdSolver,
    findorder, findorderperm, symbolicfactor, inmatrix, factor,
    triangularsolve, solve, sparspaklu, sparspaklu_refactor, ldiv, backslash
"""
from .problem import Problem
from .solver import (FP32, FP64, SparseSolver, SparseSpdSolver, backslash,
                     factor, findorder, findorderperm, inmatrix, ldiv, solve,
                     sparspaklu, sparspaklu_refactor, symbolicfactor,
                     triangularsolve)
from .symbolic import Graph, Ordering, SymbolicFactor

__all__ = [
    "Problem", "SparseSolver", "SparseSpdSolver", "findorder", "findorderperm",
    "symbolicfactor", "inmatrix", "factor", "triangularsolve", "solve",
    "sparspaklu", "sparspaklu_refactor", "ldiv", "backslash", "Graph",
    "Ordering", "SymbolicFactor", "FP32", "FP64",
]
