/** @file KttCoulomb.h
  * Tuning spaces of this library's gfx950 Coulomb kernels, for ktt programs that tune them.
  */
#pragma once

#include <cstdint>

#include <Tuner.h>

namespace ktt
{

/** Adds the CoulombSum3d tuning space (WORK_GROUP_SIZE_X/Y, POINTS_X, ROWS_PER_THREAD, ATOM_TILE,
  * INNER_UNROLL_FACTOR), its thread modifiers and constraints to a kernel whose definition has
  * the default global size (G, G*G, 1) and local size (1, 1, 1), in GlobalSizeType::CUDA. */
void AddCoulomb3dTuningSpace(Tuner& tuner, const KernelId kernel, const KernelDefinitionId definition);

/** Adds the tuning space of the slice kernels (CoulombSum2d and the CoulombSum3dIterative slice
  * kernel) for a gridX x gridY slice: the 3D parameters plus ATOM_BLOCKS (atom range split
  * over the grid's z dimension, partial sums added with atomics). The definition's default
  * global size is (gridX, gridY, 1) and local size (1, 1, 1), in GlobalSizeType::CUDA; the
  * workgroup span must tile the slice exactly, since the kernels derive the row length from
  * the launch. */
void AddCoulombSliceTuningSpace(Tuner& tuner, const KernelId kernel, const KernelDefinitionId definition, const uint64_t gridX,
    const uint64_t gridY);

/** Adds the dynamic-run space of the compensated CoulombSum3d kernel (the space above plus
  * ATOM_BLOCKS, the atom range split over workgroups along the grid's z dimension) for a grid
  * of gridSize^3 points; spans along x that waste more than an eighth of a row are excluded. */
void AddCoulombDynamicTuningSpace(Tuner& tuner, const KernelId kernel, const KernelDefinitionId definition, const uint64_t gridSize);

} // namespace ktt
