/*
 * libgwhip_h5: green-mbpt's on-disk density-fitted integrals and symmetry tables, read on the host
 * with HDF5 and handed to a libgwhip handle (V resident in HBM).
 *
 * Replaces df_integral_t's constructor, read_integrals and symmetrize
 * (src/green/mbpt/df_integral_t.h:40-187) for a kernel that keeps every k pair on the device:
 * instead of reading one chunk at a time per (k1, k2) request, every stored chunk is read once and
 * every k pair k1 >= k2 is expanded in the packed order gwhip_upload_V takes.
 *
 * Files (green-mbpt integral format >= 1.0.0):
 *   <df_path>/meta.h5        attribute __green_version__, dataset chunk_size
 *   <df_path>/VQ_<c>.h5      dataset /<c>: [chunk_size][NQ][nao][2*nao] float64 (complex interleaved),
 *                            c = first stored row of the chunk
 *   <input_h5>               /symmetry/pairs/{conj_pairs_list, trans_pairs_list, kpair_irre_list}
 *                            (all pairs stored in order when the group is absent),
 *                            /symmetry/k/{ink, bz2ibz, ibz2bz, k_sym_transform_ao, tr_conj, mesh_scaled},
 *                            /symmetry/q/{inq, bz2ibz, ibz2bz, k_sym_transform_p0, tr_conj}
 *
 * Status codes are gwhip.h's.  Not thread-safe (HDF5 is used without its thread-safe build).
 */
#ifndef GWHIP_H5_H
#define GWHIP_H5_H

#include <stdint.h>

#include "gwhip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* CheckVersion (src/green/mbpt/common_defs.h:191-212) against the format version 1.0.0:
 * 1 when `version` >= 1.0.0 (a suffix after major.minor.patch is ignored), 0 when older,
 * -1 when it does not start with major.minor.patch. */
int gwhip_h5_check_version(const char* version);

/* df_integral_t's constructor (df_integral_t.h:42-51): the version attribute of
 * <df_path>/meta.h5 (missing or older than 1.0.0 -> GWHIP_ERR_OUTDATED_INPUT) and chunk_size. */
int gwhip_h5_read_meta(const char* df_path, int64_t* chunk_size);

/* Every k pair k1 >= k2 of an nk-point mesh, packed as idx = k1(k1+1)/2 + k2:
 * V_packed [nk(nk+1)/2][NQ][nao][nao] complex, each row V(k1, k2) as read_integrals + symmetrize
 * produce it (df_integral_t.h:64-84,162-187). */
int gwhip_h5_read_packed(const char* df_path, const char* input_h5, int64_t nk, int64_t nao, int64_t NQ,
                         double* V_packed);

/* The k-symmetry tables of input_h5 (/symmetry/k) for gwhip_set_k_symmetry: *ink, bz2ibz [nk],
 * ibz2bz [nk] (first *ink entries used), U [nk][nso][nso] complex, tr_conj [nk], and the
 * fractional coordinates mesh_scaled [nk][3]. */
int gwhip_h5_read_k_symmetry(const char* input_h5, int64_t nk, int64_t nso, int64_t* ink, int64_t* bz2ibz,
                             int64_t* ibz2bz, double* U, int64_t* tr_conj, double* mesh_scaled);

/* The q-symmetry tables of input_h5 (/symmetry/q) for gwhip_set_q_symmetry: *inq, bz2ibz [nq],
 * ibz2bz [nq] (first *inq entries used), U [nq][NQ][NQ] complex (k_sym_transform_p0) and
 * tr_conj [nq]. */
int gwhip_h5_read_q_symmetry(const char* input_h5, int64_t nq, int64_t NQ, int64_t* inq, int64_t* bz2ibz,
                             int64_t* ibz2bz, double* U, int64_t* tr_conj);

/* k + q and k - q ([q][k] tables, as gwhip_set_kq_maps takes them) from the fractional k
 * coordinates: the mesh point at (k +- q) mod 1, matched to 1e-6. */
int gwhip_h5_kq_maps(const double* mesh_scaled, int64_t nk, int32_t* kplus, int32_t* kminus);

/* Everything a handle needs from the files: k +- q maps from mesh_scaled, the integrals
 * (gwhip_upload_V), and, when `with_symmetry` is non-zero, the k-symmetry tables
 * (gwhip_set_k_symmetry) and, where the file has /symmetry/q/k_sym_transform_p0, the q-symmetry
 * tables (gwhip_set_q_symmetry; switched off otherwise).  The handle's dims give nk, nao, nso
 * and NQ. */
int gwhip_h5_load(gwhip_handle* h, const char* df_path, const char* input_h5, int with_symmetry);

/* The q -> 0 extrapolation data of <df_path>/AqQ.h5 (AqQ [inq][NQ] complex, q_abs [inq], madelung;
 * mbpt_q0_utils_t's constructor, src/green/mbpt/mbpt_q0_utils_t.h:19-38) with the irreducible q of
 * input_h5 (/symmetry/q/ibz2bz; all q in order without it), passed to gwhip_set_q0_extrapolation
 * with the overlap S [ns][nk_S][nso][nso].  As in the reference, a missing AqQ.h5 switches the
 * correction off (*enabled = 0) instead of failing; an aux size other than the handle's NQ is an
 * error (GWHIP_ERR_IO). */
int gwhip_h5_load_q0(gwhip_handle* h, const char* df_path, const char* input_h5, int64_t nk_S, const double* S,
                     int* enabled);

/* Message of the last failed gwhip_h5_* call on this thread. */
const char* gwhip_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GWHIP_H5_H */
