/*
 * cmf_host.h -- host-side (CPU) map initialisation for the MI355X CaMa-Flood physics.
 *
 * These routines produce the static tables that the device physics consumes. In a
 * Fortran-hosted run the reference's own CMF_TOPO_INIT / CMF_LEVEE_INIT compute the
 * same tables and pass them through cmf_hip.h; these C entry points exist so that a
 * host without the Fortran layer (the Python host mirror, tests, benchmarks) builds
 * bit-identical tables. Each function cites the reference routine it restates.
 *
 * Conventions: all per-cell arrays are in the reference's 1-D vector (NSEQ) order,
 * 0-based. Floodplain tables are layer-planar: tab[l*n + i], l = 0..nlfp-1.
 * 2-D maps are stored with x fastest (Fortran (NX,NY) column order): map[iy*nx+ix].
 * Every function returns 0 on success and a negative code on bad arguments.
 */
#ifndef CMF_HOST_H
#define CMF_HOST_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* River-mouth / inland-end codes of the reference next-xy map (nextxy.bin). */
#define CMF_NEXT_MOUTH   (-9)
#define CMF_NEXT_INLAND  (-10)
#define CMF_NEXT_NONE    (-9999)

/* Topological 1-D ordering of the river network.
 * Restates CALC_1D_SEQ, src/cmf_ctrl_maps_mod.F90:403-525 (single region).
 * nextx/nexty: 1-based downstream cell indices (or negative codes), x fastest.
 * Outputs (capacity nx*ny): seqx/seqy 1-based grid indices of each cell,
 * next[i] = 0-based downstream cell, or the negative mouth code for mouths.
 * Returns the number of cells (NSEQALL) and writes NSEQRIV to *nseqriv. */
int cmfh_calc_1d_seq(int nx, int ny, const int32_t* nextx, const int32_t* nexty,
                     int32_t* seqx, int32_t* seqy, int32_t* next, int32_t* nseqriv);

/* Upstream lists as CSR in ascending source order (the fill order of I1UPST,
 * src/cmf_ctrl_maps_mod.F90:512-523). up_ptr has n+1 entries; up_idx capacity n. */
int cmfh_build_upstream_csr(int n, int nseqriv, const int32_t* next,
                            int32_t* up_ptr, int32_t* up_idx);

/* Channel and floodplain tables: RIVSTOMAX, RIVELV (src/cmf_ctrl_maps_mod.F90:716-722)
 * and SET_FLDSTG (src/cmf_ctrl_maps_mod.F90:969-999). lfplain=0 sets RIVSTOMAX=1e18. */
int cmfh_set_fldstg(int n, int nlfp, int lfplain,
                    const double* grarea, const double* rivlen, const double* rivwth,
                    const double* rivhgt, const double* elevtn, const double* fldhgt,
                    double* rivstomax, double* rivelv, double* fldstomax, double* fldgrd,
                    double* dfrcinc);

/* Levee tables, CMF_LEVEE_INIT section [2], src/cmf_ctrl_levee_mod.F90:148-261.
 * levhgt/levfrc are in/out (sanitised as the reference does). fldstomax/fldgrd are
 * recomputed in place exactly as the reference does at :169-181. */
int cmfh_levee_init(int n, int nlfp,
                    const double* grarea, const double* rivlen, const double* rivwth,
                    const double* rivstomax, const double* fldhgt,
                    double* fldstomax, double* fldgrd,
                    double* levhgt, double* levfrc,
                    double* bashgt, double* levdst, double* levbassto,
                    double* levtopsto, double* levfilsto);

/* Cold start: STORAGE_SEA_SURFACE, src/cmf_ctrl_vars_mod.F90:129-168.
 * Writes rivsto (FP64) and rivdph_pre. */
int cmfh_storage_sea_surface(int n, int nseqriv, const int32_t* next,
                             const double* dwnelv, const double* rivelv, const double* rivhgt,
                             const double* rivlen, const double* rivwth, const double* rivstomax,
                             double* rivsto, double* rivdph_pre);

/* ---- synthetic-network helpers (not reference restatements) ---- */

/* Priority-flood D8 drainage (Barnes et al. 2014) on a lat-lon grid that wraps in x.
 * land[k]!=0 marks land. Land cells next to sea (or beyond the y edges) seed the
 * queue as river mouths (next = CMF_NEXT_MOUTH); every other land cell drains to the
 * neighbour it was reached from. filled = depression-filled elevation, strictly
 * increasing upstream by at least eps. nextx/nexty are 1-based or negative codes. */
int cmfh_priority_flood(int nx, int ny, const float* elev, const uint8_t* land, float eps,
                        float* filled, int32_t* nextx, int32_t* nexty);

/* Upstream drainage area (sum of cell areas over the upstream tree, incl. the cell). */
int cmfh_upstream_area(int nx, int ny, const int32_t* nextx, const int32_t* nexty,
                       const double* cellarea, double* uparea);

/* Upstream accumulation on the vector network: out[r][i] = val[r][i] plus the
 * values of every cell draining to i, for nrec records of n cells. next[i] is
 * the 0-based downstream cell (< 0 at mouths). Used for the runoff climatology
 * of the spin-up surrogate (no reference counterpart). Returns 0, or -3 if the
 * network has a cycle. */
int cmfh_accumulate(int n, const int32_t* next, int nrec, const double* val, double* out);

/* Complex accumulation with a transfer factor per cell: for each of nrec records
 * (harmonics), out[i] = v[i] + sum over cells c draining to i of g[c] * out[c],
 * complex numbers as separate real / imaginary arrays [nrec][n]. With g = 1 / (1 +
 * i w tau) it propagates one harmonic of the runoff through a network of linear
 * reservoirs (residence times tau). No reference counterpart (spin-up surrogate
 * features). Returns 0, -1 on a bad index, -3 if the network has a cycle. */
int cmfh_accumulate_gain(int n, const int32_t* next, int nrec, const double* v_re, const double* v_im,
                         const double* g_re, const double* g_im, double* o_re, double* o_im);

/* Channel description for cmfh_steady_state: per-cell arrays of n cells in any
 * order (next[i] < 0 at mouths), floodplain tables layer-planar. */
typedef struct {
    int32_t n, nlfp, lfldout;
    const int32_t* next;
    const double *elevtn, *rivelv, *rivhgt, *rivwth, *rivlen, *rivman, *nxtdst, *grarea, *dwnelv;
    const double *rivstomax, *fldstomax, *fldgrd;
    double dfrcinc, pmanfld, pdstmth;
} cmfh_channel;

/* Steady state of the local-inertial river and floodplain outflow
 * (src/cmf_calc_outflw_mod.F90:60-171 with current = previous values) that
 * carries the constant outflow q[i] out of every cell: river and floodplain
 * storage, surface elevation and the river / floodplain outflow split per cell
 * (rivout, fldout may be NULL). No reference counterpart: the physical prior of
 * the spin-up surrogate. Returns 0, -1 on bad arguments, -3 on a cycle. */
int cmfh_steady_state(const cmfh_channel* c, const double* q, double* rivsto, double* fldsto, double* sfcelv,
                      double* rivout, double* fldout);

/* Linear periodic response around the steady surface elevations sfcelv: for each
 * of nh harmonics (angular frequency omega[k] in rad/s, complex local runoff
 * amplitude r[k][i] in m3/s) the complex surface-elevation amplitude h[k][i] [m]
 * of the mass balance linearised in the outflow (src/cmf_calc_outflw_mod.F90:
 * 60-171 at steady state) and the storage (stage tables), and (o_re / o_im not
 * NULL) the outflow amplitude o[k][i] [m3/s]. No reference counterpart: the
 * physical prior of the spin-up surrogate. Returns 0, -1 on bad arguments, -3 on
 * a cycle. */
int cmfh_periodic_response(const cmfh_channel* c, const double* sfcelv, int nh, const double* omega,
                           const double* r_re, const double* r_im, double* h_re, double* h_im,
                           double* o_re, double* o_im);

/* River and floodplain storage of each cell at the surface elevation sfcelv
 * (the inverse of the stage inversion of src/cmf_calc_fldstg_mod.F90:37-95 without
 * levees). No reference counterpart (surrogate). Returns 0 or -1. */
int cmfh_storage_at_level(const cmfh_channel* c, const double* sfcelv, double* rivsto, double* fldsto);

#ifdef __cplusplus
}
#endif
#endif
