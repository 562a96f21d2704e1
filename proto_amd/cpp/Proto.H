// Proto's C++ API for the MI355X hot path: include this instead of Proto's
// include/Proto.H.  Point/Box, BoxData, Stencil, Reduction, InterpStencil,
// LevelBoxData, averageDown, BoxOp, LevelOp, LevelMultigrid and
// LevelMultigridSolver compile with any C++17
// compiler and link proto_amd/libprotomi.so; forall needs hipcc (device
// lambdas) and is included when compiling with it.
#pragma once

#include "Proto_Point.H"
#include "Proto_BoxData.H"
#include "Proto_Stencil.H"
#include "Proto_InterpStencil.H"
#include "Proto_Level.H"
#include "Proto_BoxOp.H"
#include "Proto_LevelOp.H"
#include "Proto_LevelMultigrid.H"
#include "Proto_LevelMultigridSolver.H"
#ifdef __HIPCC__
#include "Proto_Forall.H"
#endif
