"""MI355X numeric phase of the Sparspak.jl sparse direct solver.

Public surface (mirrors the reference's API, Julia bangs dropped):
    Problem, SparseSolver, SparseSpdSolver,
    findorder, findorderperm, symbolicfactor, inmatrix, factor,
    triangularsolve, solve, sparspaklu, sparspaklu_refactor, ldiv, backslash,
    condest (the SPD solver's commented condition estimate, restated)
plus autotune (plan knobs chosen by timing the factorisation; no counterpart
in the reference).
"""
from .problem import Problem
from .solver import (FP32, FP64, SparseSolver, SparseSpdSolver, autotune, backslash,
                     condest, factor, findorder, findorderperm, inmatrix, ldiv, solve,
                     sparspaklu, sparspaklu_refactor, symbolicfactor,
                     triangularsolve)
from .symbolic import Graph, Ordering, SymbolicFactor

__all__ = [
    "Problem", "SparseSolver", "SparseSpdSolver", "findorder", "findorderperm",
    "symbolicfactor", "inmatrix", "factor", "triangularsolve", "solve",
    "sparspaklu", "sparspaklu_refactor", "ldiv", "backslash", "Graph",
    "Ordering", "SymbolicFactor", "FP32", "FP64", "autotune", "condest",
]
