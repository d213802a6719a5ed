/* ppf_oracle.h -- C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a host restatement of the
 * reference's PPF registration path, used by tests/, by the smoke check and by
 * bench.py's cpu_baseline leg as the checker.  The product library never links
 * it, calls it or falls back to it.
 *
 * Clouds are passed as AoS float[n*3] points and float[n*3] normals.
 * Matrices are row-major float[16].  Every function returns 0 on success.
 */
#ifndef PPF_ORACLE_H
#define PPF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- numerics probes (kernel.cu:20-30, 90-122) */
uint32_t oracle_fnv1a(const uint8_t *bytes, int n);
float oracle_quant_downf(float x, float y);
void oracle_compute_ppf(const float *p1, const float *n1, const float *p2, const float *n2,
                        float *f_raw4, float *f_quant4, float d_dist);
uint32_t oracle_feature_key(const float *f4);

/* ---- all n*n keys of one cloud, row-major [ref][partner] (ppf_kernel + ppf_hash_kernel) */
int oracle_cloud_keys(const float *pts, const float *nrm, int64_t n, int rpdf, float d_dist,
                      uint32_t *keys_out);

/* ---- parallel hash array over arbitrary 32-bit keys (parallel_hash_array.hpp:55-77)
 * key2data_out has n entries; hashkeys/counts/first have room for n entries.
 * *n_unique receives U. */
int oracle_hash_array(const uint32_t *keys, int64_t n, uint32_t *hashkeys_out,
                      uint64_t *counts_out, uint64_t *first_out, uint64_t *key2data_out,
                      int64_t *n_unique);

/* ---- alpha bin of one vote, literal trans_model_scene (kernel.cu:302-349) */
uint32_t oracle_alpha_literal(const float *m_r, const float *n_r_m, const float *m_i,
                              const float *s_r, const float *n_r_s, const float *s_i);

/* ---- voting + vote reduction (model.cu:95-171).
 * mode bit 0: 0 = literal per-vote trans_model_scene, 1 = hoisted frames;
 * mode bit 2 (| 4): keep only each reference point's best bin (largest count,
 * then smallest code) before the threshold.
 * ref_list/n_ref: optional subset of scene reference points to process (NULL =
 * all s_r with s_r % rpdf == 0).  Outputs the thresholded, ordered (code,count)
 * list; capacity cap.  stats[0..5] = lookups, hits, votes, unique codes,
 * max count, kept P.  When *n_out would exceed cap, returns 2 and sets
 * stats but not the arrays. */
int oracle_vote(const float *m_pts, const float *m_nrm, int64_t n_m,
                const float *s_pts, const float *s_nrm, int64_t n_s, float d_dist, int rpdf,
                float vote_count_threshold, int mode, const int64_t *ref_list, int64_t n_ref,
                uint64_t *codes_out, uint32_t *counts_out, int64_t cap, int64_t *n_out,
                int64_t *stats);

/* ---- poses of kept votes (trans_calc_kernel2, kernel.cu:605-645) */
int oracle_poses(const uint64_t *codes, int64_t P, const float *m_pts, const float *m_nrm,
                 const float *s_pts, const float *s_nrm, float *T_out);

/* ---- GPU-form clustering semantics (model.cu:202-244, kernel.cu:647-763),
 * weights = w[m_r]*count.  Outputs per-pose vote_counts_out, translations after
 * clustering (float3), quaternions (float4), cell hashes and the argmax. */
int oracle_cluster_spatial(const float *T, const uint64_t *codes, const uint32_t *counts,
                           const float *model_weights, int64_t P, float d_dist,
                           int use_l1_norm, int use_averaged_clusters, float *vc_out,
                           float *trans_out, float *quat_out, uint32_t *hash_out,
                           int64_t *max_idx);

/* ---- CPU greedy clustering (transformation_clustering.cpp:62-137).
 * Writes up to 3 averaged poses (row-major) and their vote sums. */
int oracle_cluster_greedy(const float *T, const uint32_t *counts, int64_t P, float trans_thresh,
                          float rot_thresh, float *poses_out, uint32_t *votes_out,
                          int64_t *n_clusters_out, int64_t *n_result);

/* ---- whole Model::ppf_lookup + pose extraction of ppf_registration
 * (ppf.cu:63-96).  Writes the final 4x4 pose and the vote stats. */
int oracle_register(const float *m_pts, const float *m_nrm, int64_t n_m, const float *s_pts,
                    const float *s_nrm, int64_t n_s, float d_dist, int rpdf,
                    float vote_count_threshold, int cpu_clustering, int use_l1_norm,
                    int use_averaged_clusters, int per_ref, float *pose_out, int64_t *stats);

/* ---- timing helper for the CPU baseline: votes for the given reference
 * points with the hoisted form, OpenMP over reference points; returns
 * lookups and votes, discards the histogram. */
int oracle_vote_throughput(const float *m_pts, const float *m_nrm, int64_t n_m,
                           const float *s_pts, const float *s_nrm, int64_t n_s, float d_dist,
                           const int64_t *ref_list, int64_t n_ref, int64_t *stats,
                           double *seconds);

float oracle_rot_angle_between(const float *A9, const float *B9);
void oracle_hrotmat2quat(const float *T16, float *q4);
int oracle_num_threads(void);
int oracle_voxel_downsample(const float *pts, const float *nrm, int64_t n, float leaf, float *out_pts,
                            float *out_nrm, int64_t *n_out);

/* ---- model handle for repeated timing calls (the hash array is built once) */
void *oracle_model_create(const float *m_pts, const float *m_nrm, int64_t n_m, float d_dist);
void oracle_model_destroy(void *h);
int oracle_vote_throughput_h(void *h, const float *m_pts, const float *m_nrm, const float *s_pts,
                             const float *s_nrm, int64_t n_s, float d_dist, const int64_t *ref_list,
                             int64_t n_ref, int64_t *stats, double *seconds);

/* ---- host evaluation of the primitives of ppf_debug_numerics (same op codes) */
int oracle_numerics(int op, const float *a, const float *b, float *out, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
