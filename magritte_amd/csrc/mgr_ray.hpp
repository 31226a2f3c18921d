// mgr_ray.hpp -- FP64 geometry arithmetic shared by the tracer and the
// radiation-field kernels.  Every expression keeps the reference's operation
// order and is compiled without floating-point contraction, so traced cell
// sequences, shifts, dZ and interpolation points are bit-identical to the
// reference's scalar code.
#pragma once

#include <cfloat>

#include "mgr_common.hpp"

namespace mgr {

// Eigen's Vector3d::dot on x86-64 (SSE2 packet of two, then the tail):
// (a.x*b.x + a.y*b.y) + a.z*b.z.
__device__ __forceinline__ double dot3(double ax, double ay, double az, double bx, double by, double bz)
{
#pragma clang fp contract(off)
    return (ax * bx + ay * by) + az * bz;
}

// One candidate of get_next, the neighbour record r.
__device__ __forceinline__ void next_candidate(const NbRec r, double ox, double oy, double oz, double nx,
                                               double ny, double nz, double Z, double &dmin, int32_t &next,
                                               double &dZ)
{
#pragma clang fp contract(off)
    const double Rx = r.x - ox;
    const double Ry = r.y - oy;
    const double Rz = r.z - oz;
    const double Z_new = dot3(Rx, Ry, Rz, nx, ny, nz);
    if (Z_new > Z) {
        const double distance_from_ray2 = dot3(Rx, Ry, Rz, Rx, Ry, Rz) - Z_new * Z_new;
        if (distance_from_ray2 < dmin) {
            dmin = distance_from_ray2;
            next = (int32_t)r.c;
            dZ = Z_new - Z;
        }
    }
}

// Geometry::get_next_general (src/Model/Geometry/geometry.tpp:328-421):
// the neighbour with Z_new > Z closest to the ray (strict <, first wins).
// (ox, oy, oz) is the origin's position; nbr holds each neighbour's index and
// position in the order of the neighbour list.
__device__ __forceinline__ int32_t get_next(const NbRec *__restrict__ nbr,
                                            const int64_t *__restrict__ nb_off, double ox, double oy,
                                            double oz, double nx, double ny, double nz, int32_t current,
                                            double &Z, double &dZ)
{
#pragma clang fp contract(off)
    double dmin = DBL_MAX;
    int32_t next = -1;
    const int64_t b = nb_off[current];
    const int deg = (int)(nb_off[current + 1] - b);
    for (int j = 0; j < deg; j++) next_candidate(nbr[b + j], ox, oy, oz, nx, ny, nz, Z, dmin, next, dZ);
    Z = Z + dZ;
    return next;
}

// get_next over the entries of the list at `list` whose bit is set in
// `cand`, in list order: the cleared entries lie behind the current cell
// along the ray by far more than the rounding of Z_new and Z, so that
// Z_new > Z fails for them.  Candidates are taken NEXT_GROUP at a time,
// their records loaded together and then tested in list order.
constexpr int NEXT_GROUP = 4;
__device__ __forceinline__ int32_t get_next_masked(const NbRec *__restrict__ list, uint64_t cand, double ox,
                                                   double oy, double oz, double nx, double ny, double nz,
                                                   double &Z, double &dZ)
{
#pragma clang fp contract(off)
    double dmin = DBL_MAX;
    int32_t next = -1;
    while (cand) {
        int j[NEXT_GROUP];
#pragma unroll
        for (int u = 0; u < NEXT_GROUP; u++) {
            j[u] = cand ? __builtin_ctzll(cand) : -1;
            cand &= cand - 1;
        }
        NbRec r[NEXT_GROUP];
#pragma unroll
        for (int u = 0; u < NEXT_GROUP; u++)
            if (j[u] >= 0) r[u] = list[j[u]];
#pragma unroll
        for (int u = 0; u < NEXT_GROUP; u++)
            if (j[u] >= 0) next_candidate(r[u], ox, oy, oz, nx, ny, nz, Z, dmin, next, dZ);
    }
    Z = Z + dZ;
    return next;
}

// Geometry::get_doppler_shift<CoMoving> (geometry.tpp:433-446)
__device__ __forceinline__ double doppler_shift(const double *__restrict__ vel, int32_t origin,
                                                int32_t current, double nx, double ny, double nz)
{
#pragma clang fp contract(off)
    const double dx = vel[3 * (int64_t)current + 0] - vel[3 * (int64_t)origin + 0];
    const double dy = vel[3 * (int64_t)current + 1] - vel[3 * (int64_t)origin + 1];
    const double dz = vel[3 * (int64_t)current + 2] - vel[3 * (int64_t)origin + 2];
    return 1.0 - dot3(dx, dy, dz, nx, ny, nz);
}

// doppler_shift from the two cells' velocities already loaded
__device__ __forceinline__ double doppler_shift_v(double vcx, double vcy, double vcz, double vox, double voy,
                                                  double voz, double nx, double ny, double nz)
{
#pragma clang fp contract(off)
    return 1.0 - dot3(vcx - vox, vcy - voy, vcz - voz, nx, ny, nz);
}

// Geometry::get_doppler_shift<Rest> (geometry.tpp:448-463): 1 - v(current).n
// with n the direction of the ray, or of its antipode for the second half of
// the rays ("up the ray").
__device__ __forceinline__ double rest_shift(const double *__restrict__ vel, int32_t current,
                                             double nx, double ny, double nz)
{
#pragma clang fp contract(off)
    return 1.0 - dot3(vel[3 * (int64_t)current + 0], vel[3 * (int64_t)current + 1],
                      vel[3 * (int64_t)current + 2], nx, ny, nz);
}

// Interpolation points of one traced step (Geometry::set_data,
// geometry.tpp:100-207).  Returns the number of projected points the step
// emits: 1 for a resolved velocity step, n_interpl for an interpolated one,
// 0 when more than 10000 points would be needed (set_data returns early).
struct StepSplit {
    int    n;       // points emitted
    int    interp;  // 1 if interpolated
    int    half;    // sub-points m < half belong to the current cell
    double dZ;      // dZ of each sub-point
    double dshift;  // shift increment of each sub-point
};

__device__ __forceinline__ StepSplit split_step(double shift_crt, double shift_nxt,
                                                double dZ_loc, double dshift_max)
{
#pragma clang fp contract(off)
    StepSplit s;
    const double dshift = shift_nxt - shift_crt;
    const double dshift_abs = fabs(dshift);
    if (dshift_abs > dshift_max) {
        const double nd = dshift_abs / dshift_max + 1;
        if (!(nd < (double)(MAX_INTERPOLATION + 1))) {   // (long)nd > 10000, or NaN
            s.n = 0; s.interp = 1; s.half = 0; s.dZ = 0.0; s.dshift = 0.0;
            return s;
        }
        const long n_interpl = (long)nd;
        s.n = (int)n_interpl;
        s.interp = 1;
        s.half = (int)(long)(0.5 * (double)n_interpl);
        s.dZ = dZ_loc / (double)n_interpl;
        s.dshift = dshift / (double)n_interpl;
    } else {
        s.n = 1; s.interp = 0; s.half = 0; s.dZ = dZ_loc; s.dshift = 0.0;
    }
    return s;
}

// shift of interpolation sub-point m: shift_crt + m * dshift_interpl
__device__ __forceinline__ double sub_shift(double shift_crt, int m, double dshift_interpl)
{
#pragma clang fp contract(off)
    return shift_crt + (double)m * dshift_interpl;
}

}  // namespace mgr
