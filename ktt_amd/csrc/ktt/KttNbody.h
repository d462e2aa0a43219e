/** @file KttNbody.h
  * Tuning space of this library's gfx950 N-body kernel (ktt_amd/kernels/Nbody.hip), the
  * counterpart of the space Examples/Nbody/Nbody.cpp:146-229 defines for the reference kernel.
  */
#pragma once

#include <cstdint>

#include <Tuner.h>

namespace ktt
{

/** Adds WORK_GROUP_SIZE_X {64,128,256}, WORK_GROUP_SIZE_Y {1,2,4} (rows splitting the sources),
 *  BODIES_PER_THREAD {1,2,4,8}, ATOM_TILE {64,128,256,512}, INNER_UNROLL_FACTOR {1,2,4,8} and
 *  USE_SOA {0,1}, the thread modifiers (local = X x Y threads; global x = bodies /
 *  (X * BODIES_PER_THREAD) groups, CUDA-style global size) and the constraints: at most 1024
 *  threads, `bodies` a multiple of X * BODIES_PER_THREAD and of ATOM_TILE * Y, and the LDS
 *  (tiles and the rows' partial sums) within 64 KiB. The definition's global size must be
 *  (bodies, 1, 1) and its local size (1, 1, 1). */
void AddNbodyTuningSpace(Tuner& tuner, const KernelId kernel, const KernelDefinitionId definition, const uint64_t bodies);

} // namespace ktt
