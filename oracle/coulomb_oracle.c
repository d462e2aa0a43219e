/*
 * coulomb_oracle.c - CPU oracle of the direct Coulomb summation path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker of the GPU library: only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it. The library never
 * calls it and has no CPU fallback.
 *
 * Quantity: V(p) = sum_i q_i / |p - r_i| on a regular grid with origin 0 and spacing h,
 * grid point (x, y, z) at (h*x, h*y, h*z), z-major layout z*G*G + y*G + x.
 *
 * Functions and what pins them (tests/test_oracle.py):
 *   oracle_coulomb3d_f32   FP32 restatement of Examples/JSON/generate-coulomb-output.c:22-31
 *                          (loop order z, y, x, atom; dx = (float)x*gs - X; q / sqrtf(r2); "+=").
 *                          Pinned bit-exactly against the output of that program compiled from
 *                          the reference sources (oracle/_ref), through the committed fixture
 *                          tests/golden/coulomb3d_kat.json (sampled values + SHA-256 of the grid).
 *   oracle_coulomb2d_f32   FP32 restatement of Examples/CoulombSum2d/CoulombSum2dReference.cl:8-20
 *                          (atoms[i].z holds dz^2; "=" store) with q / sqrtf in place of the vendor
 *                          half_rsqrt. Pinned against oracle_coulomb3d_f32 on a slice whose dz^2 is
 *                          formed as Examples/CoulombSum3dIterative/CoulombSum3dIterative.cpp:166
 *                          does, and against the FP64 evaluation.
 *   oracle_slice_dz2       dz^2 = (h*z - Z)^2 in FP32 (CoulombSum3dIterative.cpp:160-168).
 *   oracle_coulomb_points_f64
 *                          FP64 V(p) and the normaliser sum_i |q_i| / |p - r_i| at listed points,
 *                          Neumaier-compensated sums. Pinned by closed forms: a single charge
 *                          (q / r), a symmetric pair (value 0 on the bisector plane, exact dipole
 *                          sum on the axis), linearity in q, sign flip, translation invariance,
 *                          and brute force in numpy (tests/test_oracle.py).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

void oracle_coulomb3d_f32(const float* atoms, int n, float gs, int cells, int z0, int z1, float* grid)
{
    const long long plane = (long long)cells * cells;
    /* Rows (z, y) are independent: parallel over both, so that a single slice uses every thread. */
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int z = z0; z < z1; z++)
    {
        for (int y = 0; y < cells; y++)
        {
            for (int x = 0; x < cells; x++)
            {
                float* out = &grid[(long long)z * plane + (long long)y * cells + x];
                /* The same additions in the same order as `*out += e` per atom, kept in a register
                 * (the grid cannot alias the atoms, but the compiler cannot know it). */
                float acc = *out;
                for (int at = 0; at < n; at++)
                {
                    const float dx = (float)x * gs - atoms[at * 4 + 0];
                    const float dy = (float)y * gs - atoms[at * 4 + 1];
                    const float dz = (float)z * gs - atoms[at * 4 + 2];
                    const float e = atoms[at * 4 + 3] / sqrtf(dx * dx + dy * dy + dz * dz);
                    acc += e;
                }
                *out = acc;
            }
        }
    }
}

void oracle_coulomb2d_f32(const float* atoms, int n, float gs, int gx, int gy, float* grid)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < gy; y++)
    {
        for (int x = 0; x < gx; x++)
        {
            const float coordX = gs * (float)x;
            const float coordY = gs * (float)y;
            float energy = 0.0f;
            for (int i = 0; i < n; i++)
            {
                const float dx = coordX - atoms[4 * i + 0];
                const float dy = coordY - atoms[4 * i + 1];
                energy += atoms[4 * i + 3] / sqrtf(dx * dx + dy * dy + atoms[4 * i + 2]);
            }
            grid[(long long)y * gx + x] = energy;
        }
    }
}

void oracle_slice_dz2(const float* atoms, int n, float gs, int z_index, float* out)
{
    const float z = gs * (float)z_index;
    for (int i = 0; i < n; i++)
    {
        out[4 * i + 0] = atoms[4 * i + 0];
        out[4 * i + 1] = atoms[4 * i + 1];
        out[4 * i + 2] = (z - atoms[4 * i + 2]) * (z - atoms[4 * i + 2]);
        out[4 * i + 3] = atoms[4 * i + 3];
    }
}

/* Neumaier summation of (sum, c) += term. */
static inline void add_compensated(double* sum, double* c, const double term)
{
    const double t = *sum + term;
    if (fabs(*sum) >= fabs(term))
    {
        *c += (*sum - t) + term;
    }
    else
    {
        *c += (term - t) + *sum;
    }
    *sum = t;
}

/* points: m triples (x, y, z) of grid indices; h as the FP32 spacing the GPU uses, widened. */
void oracle_coulomb_points_f64(const float* atoms, int n, float gs, const int32_t* points, int m, double* values,
    double* normalisers)
{
    const double h = (double)gs;
#pragma omp parallel for schedule(dynamic, 16)
    for (int p = 0; p < m; p++)
    {
        const double px = h * (double)points[3 * p + 0];
        const double py = h * (double)points[3 * p + 1];
        const double pz = h * (double)points[3 * p + 2];
        double sum = 0.0, c = 0.0, abs_sum = 0.0, abs_c = 0.0;
        for (int i = 0; i < n; i++)
        {
            const double dx = px - (double)atoms[4 * i + 0];
            const double dy = py - (double)atoms[4 * i + 1];
            const double dz = pz - (double)atoms[4 * i + 2];
            const double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            const double q = (double)atoms[4 * i + 3];
            add_compensated(&sum, &c, q * inv);
            add_compensated(&abs_sum, &abs_c, fabs(q) * inv);
        }
        values[p] = sum + c;
        normalisers[p] = abs_sum + abs_c;
    }
}

/* FP64 evaluation on a slice whose atoms carry dz^2 in .z (the 2D semantics). */
void oracle_coulomb2d_points_f64(const float* atoms, int n, float gs, const int32_t* points, int m, double* values,
    double* normalisers)
{
    const double h = (double)gs;
#pragma omp parallel for schedule(dynamic, 16)
    for (int p = 0; p < m; p++)
    {
        const double px = h * (double)points[2 * p + 0];
        const double py = h * (double)points[2 * p + 1];
        double sum = 0.0, c = 0.0, abs_sum = 0.0, abs_c = 0.0;
        for (int i = 0; i < n; i++)
        {
            const double dx = px - (double)atoms[4 * i + 0];
            const double dy = py - (double)atoms[4 * i + 1];
            const double inv = 1.0 / sqrt(dx * dx + dy * dy + (double)atoms[4 * i + 2]);
            const double q = (double)atoms[4 * i + 3];
            add_compensated(&sum, &c, q * inv);
            add_compensated(&abs_sum, &abs_c, fabs(q) * inv);
        }
        values[p] = sum + c;
        normalisers[p] = abs_sum + abs_c;
    }
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
