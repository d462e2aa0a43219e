/* CPU oracle of the N-body integration step (test infrastructure, never the product).
 *
 * The step is the one of Examples/Nbody/NbodyReference.cu:40-75 and Nbody.cu:271-410:
 *   a_i  = sum_j m_j (p_j - p_i) / (|p_j - p_i|^2 + eps^2)^(3/2)   (j = i contributes 0)
 *   p_i' = p_i + dt v_i + damping dt^2 a_i,   v_i' = v_i + dt a_i
 * with positions and velocities as float4 (x, y, z, mass) and (vx, vy, vz, 0); the mass and
 * the velocity's w pass through.
 *
 *   oracle_nbody_f32   FP32 restatement in the reference kernel's order: j ascending,
 *                      1/sqrtf for the reference's rsqrtf, float4 arithmetic of
 *                      NbodyReference.cu:60-70 (the w lane included).
 *   oracle_nbody_f64   FP64 with compensated (Neumaier) sums; also returns the accelerations.
 *
 * Pinned by (tests/test_oracle.py): two bodies (the closed form along their axis), a body at
 * the centre of a cube of equal masses (zero acceleration by symmetry), momentum
 * conservation (sum_i m_i a_i = 0 for the antisymmetric pair forces), numpy brute force on
 * small inputs, the FP32 restatement within its rounding of the FP64 one, and planted mistakes
 * (a sign, the softening, the power of r) that must fail those checks.
 */
#include <math.h>
#include <stdint.h>

static inline void neumaier(double* sum, double* c, const double term)
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

void oracle_nbody_f32(const float* pos, const float* vel, int n, float dt, float damping, float soft2, float* pos_out,
                      float* vel_out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
    {
        const float px = pos[4 * i], py = pos[4 * i + 1], pz = pos[4 * i + 2], pw = pos[4 * i + 3];
        float ax = 0.0f, ay = 0.0f, az = 0.0f, aw = 0.0f;
        for (int j = 0; j < n; ++j)
        {
            const float dx = pos[4 * j] - px;
            const float dy = pos[4 * j + 1] - py;
            const float dz = pos[4 * j + 2] - pz;
            const float dw = pos[4 * j + 3] - pw;
            const float invr = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz + soft2);
            const float f = pos[4 * j + 3] * invr * invr * invr;
            ax += f * dx;
            ay += f * dy;
            az += f * dz;
            aw += f * dw;
        }
        /* p += dt*v + damping*dt*dt*a and v += dt*a with dt = (dt, dt, dt, 0). */
        const float dtdt = damping * dt * dt;
        pos_out[4 * i] = px + (dt * vel[4 * i] + dtdt * ax);
        pos_out[4 * i + 1] = py + (dt * vel[4 * i + 1] + dtdt * ay);
        pos_out[4 * i + 2] = pz + (dt * vel[4 * i + 2] + dtdt * az);
        pos_out[4 * i + 3] = pw + (0.0f * vel[4 * i + 3] + (damping * 0.0f * 0.0f) * aw);
        vel_out[4 * i] = vel[4 * i] + dt * ax;
        vel_out[4 * i + 1] = vel[4 * i + 1] + dt * ay;
        vel_out[4 * i + 2] = vel[4 * i + 2] + dt * az;
        vel_out[4 * i + 3] = vel[4 * i + 3] + 0.0f * aw;
    }
}

void oracle_nbody_f64(const float* pos, const float* vel, int n, float dt, float damping, float soft2, double* pos_out,
                      double* vel_out, double* acc_out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
    {
        const double px = pos[4 * i], py = pos[4 * i + 1], pz = pos[4 * i + 2];
        double sx = 0.0, sy = 0.0, sz = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
        for (int j = 0; j < n; ++j)
        {
            const double dx = (double)pos[4 * j] - px;
            const double dy = (double)pos[4 * j + 1] - py;
            const double dz = (double)pos[4 * j + 2] - pz;
            const double r2 = dx * dx + dy * dy + dz * dz + (double)soft2;
            const double f = (double)pos[4 * j + 3] / (r2 * sqrt(r2));
            neumaier(&sx, &cx, f * dx);
            neumaier(&sy, &cy, f * dy);
            neumaier(&sz, &cz, f * dz);
        }
        const double a[3] = {sx + cx, sy + cy, sz + cz};
        for (int k = 0; k < 3; ++k)
        {
            const double v = vel[4 * i + k];
            acc_out[3 * i + k] = a[k];
            pos_out[4 * i + k] = (double)pos[4 * i + k] + (double)dt * v + (double)damping * (double)dt * (double)dt * a[k];
            vel_out[4 * i + k] = v + (double)dt * a[k];
        }
        pos_out[4 * i + 3] = pos[4 * i + 3];
        vel_out[4 * i + 3] = vel[4 * i + 3];
    }
}
