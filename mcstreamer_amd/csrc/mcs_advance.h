// mcs_advance.h -- the per-particle body of JuMC.advance! shared by the
// single-population kernel (mcs_step.hip) and the batched swarm kernel
// (mcs_swarm.hip).
#pragma once

#include <type_traits>

#include "mcs_device.h"

// The photoemission sink of a step without photons: never called (tables with a
// photoemission row are refused before such a step starts).
struct NoPhoto {
    __device__ void operator()(int, const double *, uint8_t, int64_t) const {}
};

// track1 (MCStreamer/src/gridfields.jl:203-213): the NGP cell of x (flat index
// into the ghosted cell-centred array) and the amount val/dV it receives; false
// when x is outside the mesh.  FIELD 2 is the axisymmetric grid (dV = 2 pi rc dr
// dz of the cell's radial index), otherwise the 3D mesh (dV = h^3).  Every mesh
// array has fewer than 2^31 elements (checked at creation).
template <int FIELD>
__device__ __forceinline__ bool track_cell(const MeshView &M, const double x[3], double val, int32_t &cell,
                                           double &q)
{
    if (FIELD == 2) {
        int64_t i;
        const int64_t c = cyl_cell_dev(M.C, x, i);
        if (c < 0) return false;
        cell = (int32_t)c;
        q = val / gld(M.C.dv + (i - 1));
        return true;
    }
    int64_t c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double f = floor(x[d] / M.h);
        if (!(f >= 0.0 && f < (double)M.n[d])) return false;
        c[d] = (int64_t)f;
    }
    const double dV = (M.h * M.h) * M.h;
    cell = (int32_t)((c[0] + 1) + (int64_t)(M.n[0] + 2) * ((c[1] + 1) + (int64_t)(M.n[1] + 2) * (c[2] + 1)));
    q = val / dV;
    return true;
}

// The field at x: uniform (FIELD 0), the 3D mesh's faces (1) or the
// axisymmetric grid's (2).
template <int FIELD>
__device__ __forceinline__ void field_at(const MeshView &M, const double x[3], const double *euni, double e[3],
                                         int32_t &flags)
{
    if (FIELD == 2) {
        efield_cyl_dev(M.C, x, e, flags);
    } else if (FIELD == 1) {
        efield_mesh_dev(M, x, e, flags);
    } else {
        e[0] = euni[0];
        e[1] = euni[1];
        e[2] = euni[2];
    }
}

// One particle through one time step dt (JuMC/src/mixed_population.jl:31-51):
// free flights to the next (null) collision, the Yoshida push of
// electron.jl:51-62 with the field evaluated three times per flight, the
// sampler of collisions.jl:159-192 and the outcomes of electron.jl:108-139.
// `birth(x, v, s, active)` receives each ionization child (weight w), `event(j)`
// each selected table row, and with TRACK `track(x, +-w)` each ion-charge deposit
// of the CollisionTracker (gridfields.jl:203-218: +w at an ionization, -w at an
// attachment, at the particle's position).  With a `photo` sink, a PhotoEmission
// row (electron.jl:141-144) draws its photon count Poisson((w / multiplier) /
// photon_weight) from the particle's stream and passes (row, x, active, count) to
// `photo`; the electron is unchanged.  A particle removed by attachment keeps
// flying until dt is used up (the loop does not test `active`).
template <int FIELD_MESH, int TRACK, typename Birth, typename Event, typename Track, typename Photo = NoPhoto>
__device__ __forceinline__ void advance_one(double x[3], double v[3], const double w, double &s,
                                            uint8_t &act, mcs_stream &rs, const double dt,
                                            const TableView &T, const MeshView &M,
                                            const double *euni, int32_t &flags, int &flights,
                                            int &deaths, Birth &&birth, Event &&event, Track &&track,
                                            Photo &&photo = Photo())
{
    constexpr bool PHOTO = !std::is_same<typename std::decay<Photo>::type, NoPhoto>::value;
    const double maxrate = T.maxrate;
    double trem = dt;
    const double c1 = MCS_YOSHIDA_C1, c2 = MCS_YOSHIDA_C2;
    while (trem > 0.0) {
        double tnext = s / maxrate;
        double tt;
        bool collides;
        if (trem > tnext) {
            tt = tnext;
            collides = true;
        } else {
            tt = trem;
            collides = false;
            s = s - tt * maxrate;
        }
        // advance_free_yoshida
        {
            double e[3];
            double k1 = c1 * tt, k2 = c2 * tt;
            double f1 = (MCS_YOSHIDA_D1 * tt) * MCS_QM;
            double f2 = (MCS_YOSHIDA_D2 * tt) * MCS_QM;
#pragma unroll
            for (int d = 0; d < 3; ++d) x[d] = x[d] + k1 * v[d];
            field_at<FIELD_MESH>(M, x, euni, e, flags);
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] = v[d] + f1 * e[d];
#pragma unroll
            for (int d = 0; d < 3; ++d) x[d] = x[d] + k2 * v[d];
            field_at<FIELD_MESH>(M, x, euni, e, flags);
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] = v[d] + f2 * e[d];
#pragma unroll
            for (int d = 0; d < 3; ++d) x[d] = x[d] + k2 * v[d];
            field_at<FIELD_MESH>(M, x, euni, e, flags);
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] = v[d] + f1 * e[d];
#pragma unroll
            for (int d = 0; d < 3; ++d) x[d] = x[d] + k1 * v[d];
        }
        flights += 1;
        if (collides) {
            double eng = energy_dev(v[0], v[1], v[2]);
            int64_t k;
            double wg;
            indweight_dev(T.energy, T.ne, T.egrid, T.step, T.inv_dlog, eng, k, wg, flags);
            double xi = mcs_u01(&rs) * maxrate;
            const double *r0 = T.rate + k * T.nproc;
            const double *r1 = r0 + T.nproc;
            // the selected row first (do_one_collision!), then its outcome
            int sel = -1;
            for (int j = 0; j < T.nproc; ++j) {
                double nu = wg * gld(r0 + j) + (1.0 - wg) * gld(r1 + j);
                if (nu > xi) {
                    sel = j;
                    break;
                }
                xi = xi - nu;
            }
            if (sel >= 0) {
                event(sel);
                int kind = gld(T.kind + sel);
                if (kind == MCS_PROC_ELASTIC) {
                    double mr = gld(T.param + sel);
                    double c = mr / (1.0 + mr);
                    double vcm0 = c * v[0], vcm1 = c * v[1], vcm2 = c * v[2];
                    double d0 = v[0] - vcm0, d1 = v[1] - vcm1, d2 = v[2] - vcm2;
                    double nrm = sqrt((d0 * d0 + d1 * d1) + d2 * d2);
                    double om[3];
                    mcs_randsphere(&rs, om);
                    v[0] = om[0] * nrm + vcm0;
                    v[1] = om[1] * nrm + vcm1;
                    v[2] = om[2] * nrm + vcm2;
                } else if (kind == MCS_PROC_EXCITATION) {
                    double e1 = eng - gld(T.param + sel);
                    if (e1 < 0.0) e1 = 0.0;
                    double vabs = sqrt((2.0 * e1) / MCS_ELECTRON_MASS);
                    double om[3];
                    mcs_randsphere(&rs, om);
                    v[0] = om[0] * vabs;
                    v[1] = om[1] * vabs;
                    v[2] = om[2] * vabs;
                } else if (kind == MCS_PROC_IONIZATION) {
                    double e1 = eng - gld(T.param + sel);
                    if (e1 < 0.0) e1 = 0.0;
                    e1 = e1 / 2.0;
                    double vabs = sqrt((2.0 * e1) / MCS_ELECTRON_MASS);
                    double om1[3], om2[3];
                    mcs_randsphere(&rs, om1);
                    mcs_randsphere(&rs, om2);
                    v[0] = om1[0] * vabs;
                    v[1] = om1[1] * vabs;
                    v[2] = om1[2] * vabs;
                    double sc = mcs_nextcoll(&rs);
                    if (TRACK) track(x, w);
                    double vc[3] = {om2[0] * vabs, om2[1] * vabs, om2[2] * vabs};
                    birth(x, vc, sc, act);
                } else if (kind == MCS_PROC_ATTACHMENT) {
                    if (TRACK) track(x, -w);
                    if (act) deaths += 1;
                    act = 0;
                } else if (PHOTO && kind == MCS_PROC_PHOTOEMISSION) {
                    const double *ph = T.photo + 4 * sel;
                    const int64_t np = mcs_poisson(&rs, (w / gld(ph + 2)) / gld(ph + 3));
                    photo(sel, x, act, np);
                }
            }
            s = mcs_nextcoll(&rs);
        }
        trem = trem - tt;
    }
}
