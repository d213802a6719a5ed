/* ppf_amd.h -- C ABI of the MI355X PPF registration library (libppf_amd.so).
 *
 * Drop-in boundary for the point-pair-feature registration path of
 * nicolasavru/ppf-registration-spatial-hashing.  Plain C types only: host
 * pointers and sizes in, opaque handles and status codes out.  Every function
 * returns PPF_OK (0) on success, a negative/positive status otherwise, with a
 * thread-local message in ppf_last_error().  Host inputs are copied during the
 * call; nothing is borrowed after return.  Nothing calls exit().
 *
 * Clouds are AoS float[n*3] xyz points and float[n*3] normals (the fields of
 * pcl::PointNormal).  Poses are row-major float[16].
 *
 * A context owns one HIP stream on one device; a context and the objects made
 * from it must not be used from two threads at once.
 */
#ifndef PPF_AMD_H
#define PPF_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define PPF_OK 0
#define PPF_ERR_ARG 1        /* invalid argument (sizes, NULL pointers, limits) */
#define PPF_ERR_HIP 2        /* a HIP runtime call failed */
#define PPF_ERR_NOMEM 3      /* device allocation or HBM budget exhausted */
#define PPF_ERR_NO_VOTES 4   /* no scene pair matched the model: no pose */
#define PPF_ERR_STATE 5      /* wrong object state (e.g. result not computed) */
#define PPF_ERR_TIMEOUT 6    /* ppf_server_wait: the request is not done yet */
#define PPF_ERR_CANCELLED 7  /* the request was cancelled or the server stopped */

/* The reference encodes votes as [s_r:32 | m_r:26 | alpha:6] (include/model.h:61-63)
 * and indexes model pairs m_r*n_m + m_i; this build keeps pair indices in 32 bits. */
#define PPF_MAX_MODEL_POINTS 65535

typedef struct ppf_ctx_s *ppf_ctx;
typedef struct ppf_model_s *ppf_model;
typedef struct ppf_scene_s *ppf_scene;
typedef struct ppf_result_s *ppf_result;

/* Last error message of the calling thread ("" if none).
 * Replaces HANDLE_ERROR's fprintf+exit (include/impl/util.hpp:18-26). */
const char *ppf_last_error(void);

/* ------------------------------------------------------------------ context
 * Device selection as ppf_registration does it: min(numDevices-1, dev)
 * (src/cuda/ppf.cu:37-49).  hbm_budget_bytes = 0 means no budget. */
int ppf_ctx_create(int device, uint64_t hbm_budget_bytes, ppf_ctx *out);
int ppf_ctx_destroy(ppf_ctx ctx);
/* Bytes of device memory currently held by objects of this context. */
uint64_t ppf_ctx_device_bytes(ppf_ctx ctx);

/* Progress reports of long lookups (the reference's DEBUG stage timers,
 * src/cuda/scene.cu:72-98 and model.cu:271-305, as a callback).  With a
 * callback set, ppf_lookup on this context votes in launches of
 * rows_per_launch reference rows (0: all rows in one launch) and calls
 * fn(user, stage, done, total) after each launch ("vote": rows voted of the
 * scene's reference rows) and after each later stage ("select", "poses",
 * "cluster": done = total = kept votes).  ppf_batch_lookup likewise votes
 * its jobs' rows in launches of about rows_per_launch rows and reports
 * ("vote": rows voted of all jobs' reference rows) after each.  Results do
 * not depend on it.  fn runs on the calling thread; NULL turns reports off. */
typedef void (*ppf_progress_fn)(void *user, const char *stage, int64_t done, int64_t total);
int ppf_ctx_set_progress(ppf_ctx ctx, ppf_progress_fn fn, void *user, int64_t rows_per_launch);

/* -------------------------------------------------------------------- model
 * Model::Model(cloud, d_dist, ...) (include/model.h:17-19, src/cuda/model.cu:43-82):
 * PPF of every ordered model pair, quantized and FNV-hashed, then the sort-based
 * parallel hash array (include/impl/parallel_hash_array.hpp:55-77).
 * The vote-side parameters of the reference constructor are per lookup here
 * (ppf_lookup_params). */
int ppf_model_build(ppf_ctx ctx, const float *points, const float *normals, int64_t n,
                    float d_dist, ppf_model *out);
int ppf_model_destroy(ppf_model model);

typedef struct {
  int64_t n_points;
  float d_dist;
  int64_t n_pairs;       /* n*n, diagonal included as in the reference */
  int64_t n_unique_keys; /* U, key 0 included when present */
  int32_t tile_points;   /* model points per vote accumulator tile */
  int32_t n_tiles;
  uint64_t device_bytes;
} ppf_model_info_t;
int ppf_model_info(ppf_model model, ppf_model_info_t *info);

/* Copy the parallel hash array to the host: ParallelHashArray::GetHashkeys,
 * GetCounts, GetFirstHashkeyIndices, GetHashkeyToDataMap
 * (include/impl/parallel_hash_array.hpp:94-112).  Arrays sized U, U, U, n*n;
 * any pointer may be NULL. */
int ppf_model_hash_array(ppf_model model, uint32_t *hashkeys, uint32_t *counts, uint32_t *first,
                         uint32_t *key2ppf);

/* Save a built model (its device arrays, with a versioned header and a
 * checksum) and load it on a context, for warm starts without rebuilding the
 * hash array (the reference only dumps raw device vectors,
 * include/impl/util.hpp:55-91).  Loading checks the version, the header's
 * consistency, that the model's tiles fit this device's LDS, and the
 * checksum. */
int ppf_model_save(ppf_model model, const char *path);
int ppf_model_load(ppf_ctx ctx, const char *path, ppf_model *out);

/* Model::SetModelPointVoteWeights (src/cuda/model.cu:91-93): w[m_r] scales the
 * vote count of every kept vote of model point m_r before GPU clustering
 * (vote_weight_kernel, src/cuda/kernel.cu:766-782).  Default all 1. */
int ppf_model_set_weights(ppf_model model, const float *weights, int64_t n);

/* -------------------------------------------------------------------- scene
 * Scene::Scene(cloud, d_dist, ref_point_downsample_factor) (include/scene.h:15-16,
 * src/cuda/scene.cu:24-55).  Scene features depend on d_dist, which belongs to
 * the model, so they are computed inside ppf_lookup; the scene object holds the
 * points and the local frames of its reference points (idx % rpdf == 0). */
int ppf_scene_upload(ppf_ctx ctx, const float *points, const float *normals, int64_t n,
                     uint32_t ref_point_downsample_factor, ppf_scene *out);
int ppf_scene_destroy(ppf_scene scene);

/* All n*n scene keys for the given d_dist, row-major [s_r][s_i]; 0 for rows
 * that are not reference points and for the diagonal: Scene::getHashKeys after
 * ppf_kernel + ppf_hash_kernel (src/cuda/scene.cu:64-99, kernel.cu:404-477).
 * For inspection of small clouds only (n <= 8192). */
int ppf_scene_keys(ppf_scene scene, float d_dist, uint32_t *keys_out);

/* ------------------------------------------------------------------- lookup
 * Model::ppf_lookup(Scene*) (include/model.h:35, src/cuda/model.cu:269-306):
 * voting, vote threshold, poses, weighted counts, and clustering. */
typedef struct {
  float vote_count_threshold;  /* keep votes with count > thr * max (model.cu:164-170) */
  int32_t cpu_clustering;      /* 1: greedy host clustering (transformation_clustering.cpp) */
  int32_t use_l1_norm;         /* 1: skip the translation test (kernel.cu:740-746) */
  int32_t use_averaged_clusters;
  int32_t exact_votes;         /* 1: evaluate every alpha bin with the full float path */
  uint32_t candidate_capacity; /* initial (code, count) buffer capacity; 0 = automatic (4,096
                                  per reference point, at least 8M).  A smaller buffer than
                                  the vote pass emits costs one rerun of it, never results */
  int32_t per_reference_argmax; /* 1: keep only each reference point's best (model point,
                                   alpha) bin -- larger count, then smaller code -- before the
                                   threshold: the score pass that include/kernel.h:78-82
                                   declares (ppf_score_kernel) and never defines */
} ppf_lookup_params;

int ppf_lookup(ppf_model model, ppf_scene scene, const ppf_lookup_params *params,
               ppf_result *out);
int ppf_result_destroy(ppf_result result);


typedef struct {
  int32_t status;          /* PPF_OK or PPF_ERR_NO_VOTES */
  int64_t lookups;         /* scene pairs looked up: R*(n_s-1) */
  int64_t votes;           /* nonunique votes: matched (scene pair, model pair) combinations */
  int64_t max_count;       /* largest (s_r, m_r, alpha) count */
  int64_t n_candidates;    /* (code,count) pairs emitted before the final threshold */
  int64_t n_kept;          /* P: votes kept after the threshold */
  int64_t max_idx;         /* GPU clustering argmax (model->max_idx), -1 for CPU */
  int64_t n_clusters;      /* CPU clustering: number of clusters */
  double vote_ms;          /* device time of the vote kernel (HIP events) */
  double total_ms;         /* device time of the whole lookup */
  double select_ms;        /* threshold + ordering of the kept votes */
  double pose_ms;          /* one pose per kept vote */
  double cluster_ms;       /* clustering (GPU or host) and pose extraction */
  int64_t cluster_tests;   /* GPU clustering: (pose, neighbour pose) pairs tested */
  int64_t vote_rows_ungrouped;  /* plain lookups: reference rows whose hits spanned more model
                                   buckets than the vote kernel groups, voted hit by hit */
  int64_t cluster_mixed_chunks; /* GPU clustering: pose chunks spanning several cells that
                                   share a spatial hash (FNV collision) */
  double cluster_kernel_ms;     /* GPU clustering: device time of the pair-test kernel alone */
} ppf_result_info_t;
int ppf_result_info(ppf_result result, ppf_result_info_t *info);

/* Allocation-free Model::ppf_lookup.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same stages and the
 * same results as ppf_lookup + ppf_result_pose/ppf_result_info, but every
 * buffer it needs beyond the model, the scene and the caller's outputs is a
 * piece of one static device array of PPF_NOALLOC_SCRATCH_BYTES per device
 * (held until the process ends); it calls no hipMalloc/hipFree.  Limits: the
 * scene has at most PPF_NOALLOC_MAX_SCENE_POINTS points (PPF_ERR_ARG above),
 * and a lookup whose buffers do not fit the array (many kept votes) returns
 * PPF_ERR_NOMEM with a message.  All calls on one GPU share its array: calls
 * that overlap on one GPU (from several threads or contexts) race.  Returns
 * PPF_OK, or PPF_ERR_NO_VOTES with a zero pose, like ppf_result_pose. */
#define PPF_NOALLOC_SCRATCH_BYTES (256ull << 20)
#define PPF_NOALLOC_MAX_SCENE_POINTS 8192
int ppf_lookup_noalloc(ppf_model model, ppf_scene scene, const ppf_lookup_params *params,
                       float *pose_out, ppf_result_info_t *info_out);

/* Kept votes in reference order (count descending, code ascending):
 * Model::votes / Model::voteCounts after ComputeUniqueVotes (model.cu:148-170). */
int ppf_result_votes(ppf_result result, uint64_t *codes, uint32_t *counts);
/* One row-major 4x4 per kept vote: Model::getTransformations (model.cu:191-200,308-310). */
int ppf_result_transforms(ppf_result result, float *transforms);
/* GPU clustering outputs per kept vote: vote_counts_out, transformation_trans,
 * transformation_rots (model.cu:202-244).  Any pointer may be NULL. */
int ppf_result_clusters(ppf_result result, float *vote_counts_out, float *translations,
                        float *quaternions);
/* CPU clustering outputs: Model::cpu_transformations, the (at most three) best
 * clusters of clusterPoses, most votes first (transformation_clustering.cpp:
 * 96-122): poses_out [3][16] averaged poses, votes_out [3] cluster vote
 * sums, *n_out their number.  poses_out and votes_out may be NULL.
 * PPF_ERR_STATE when the lookup used GPU clustering. */
int ppf_result_cpu_clusters(ppf_result result, float *poses_out, uint32_t *votes_out,
                            int32_t *n_out);
/* Final pose of the (scene, model) pair as ppf_registration extracts it
 * (src/cuda/ppf.cu:74-93). */
int ppf_result_pose(ppf_result result, float *pose16);

/* ----------------------------------------------------------- registration
 * ppf_registration (include/ppf.h:9-15, src/cuda/ppf.cu:29-106): for every
 * scene and every model, the pose of the model in the scene.
 * scene_points[i]/scene_normals[i] have scene_sizes[i] points; same for models.
 * poses_out: [n_scenes][n_models][16]; status_out (optional): [n_scenes][n_models].
 * model_weights: optional per-model arrays (NULL entries allowed). */
int ppf_registration(ppf_ctx ctx, int32_t n_scenes, const float *const *scene_points,
                     const float *const *scene_normals, const int64_t *scene_sizes,
                     int32_t n_models, const float *const *model_points,
                     const float *const *model_normals, const int64_t *model_sizes,
                     const float *model_d_dists, uint32_t ref_point_downsample_factor,
                     float vote_count_threshold, int32_t cpu_clustering, int32_t use_l1_norm,
                     int32_t use_averaged_clusters, const float *const *model_weights,
                     float *poses_out, int32_t *status_out);

/* ------------------------------------------------------------ pre-processing
 * Voxel-grid downsampling as the alignment driver applies it before
 * registration (pcl::VoxelGrid with leaf (l, l, l), src/alignment.cpp:78-87,
 * 259-288: the scene with scene_leaf_size, each model with its d_dist).
 * Points with non-finite x, y or z are dropped; each occupied voxel becomes
 * the average of its points (positions and normals), voxels in ascending
 * voxel index; a leaf too small for the cloud's extent (more than INT32_MAX
 * voxels) returns the input unchanged, as PCL does.  points_out/normals_out
 * (float[n*3] suffices) may be NULL to get only the count. */
int ppf_voxel_downsample(ppf_ctx ctx, const float *points, const float *normals, int64_t n, float leaf,
                         float *points_out, float *normals_out, int64_t *n_out);
/* max - min per axis over the finite points (pcl::getMinMax3D,
 * src/alignment.cpp:246-257); the driver's d_dist is tau_d * the largest. */
int ppf_bbox_extent(const float *points, int64_t n, float extent_out[3]);

/* ------------------------------------------------ files and validation (host)
 * PLY point clouds with normals as the driver loads them
 * (pcl::io::loadPLYFile<PointNormal>, src/alignment.cpp:207-244): the vertex
 * element's x y z and nx ny nz (or normal_x ...), ascii or binary, any
 * numeric property types; other properties and elements are skipped and
 * missing normals read as 0.  Call with NULL buffers to get *n_out first. */
int ppf_ply_read(const char *path, float *points, float *normals, int64_t capacity, int64_t *n_out);
/* Writes x y z (and nx ny nz when normals is not NULL) as float properties. */
int ppf_ply_write(const char *path, const float *points, const float *normals, int64_t n, int32_t binary);
/* 16 whitespace-separated numbers, row-major (include/impl/util.hpp:95-104). */
int ppf_xf_read(const char *path, float pose16[16]);
/* ht_dist (src/cuda/linalg.cu:9-20): out2[0] = |t_a - t_b|, out2[1] = angle
 * of rot_a^-1 rot_b in radians; the driver's validation accepts a pose when
 * out2[0] < validation_translation_threshold * diameter and out2[1] <
 * validation_rotation_threshold (src/alignment.cpp:317-323). */
int ppf_ht_dist(const float a16[16], const float b16[16], float out2[2]);

/* ----------------------------------------------------------------- batched
 * Many (model, scene) pairs in one call: the batched form of the loop in
 * ppf_registration (src/cuda/ppf.cu:63-98) over device-resident models.
 * All jobs are voted in one persistent launch (work items are (job, reference
 * point), longest jobs first).  The stages after voting then run once for
 * all jobs -- one threshold/order sort, one pose launch, one clustering
 * launch over every job's pose chunks (heaviest first), one argmax -- or,
 * with CPU clustering or when the batch's sort keys would exceed 64 bits,
 * job by job concurrently on the context's lanes.  results_out[j] is the
 * result of (models[j], scenes[j]) and is bit-identical to
 * ppf_lookup(models[j], scenes[j], params); the results of a batch finished
 * together share one device allocation, freed with the last of them.
 * A model or scene may appear in several jobs.  info may be NULL. */
#define PPF_MAX_BATCH_JOBS 65536
typedef struct {
  int64_t n_jobs;
  int64_t lookups;       /* sum over jobs */
  int64_t votes;         /* sum over jobs */
  int32_t vote_launches; /* 1, or more with progress reports or when a job's candidate buffer overflowed */
  int32_t lanes;         /* lanes that finished the jobs (0: finished together) */
  double vote_ms;        /* device time of the vote launches */
  double finish_ms;      /* wall time of threshold/order, poses and clustering */
  double total_ms;       /* wall time of the call */
  int32_t fused;         /* 1: the stages after voting ran once for all jobs */
  int32_t vote_graph;    /* 1: the vote stage ran as captured HIP graphs, replayed
                            while the batch's launch signature stays the same */
  double select_ms;      /* fused: device time of threshold/order of all jobs */
  double pose_ms;        /* fused: device time of the poses of all jobs */
  double cluster_ms;     /* fused: device time of clustering and final poses */
  double cluster_kernel_ms; /* fused: the pair-test kernel alone */
} ppf_batch_info_t;
int ppf_batch_lookup(ppf_ctx ctx, int32_t n_jobs, const ppf_model *models, const ppf_scene *scenes,
                     const ppf_lookup_params *params, ppf_result *results_out,
                     ppf_batch_info_t *info);

/* Autotuner of the batched form (no counterpart in the reference, whose
 * launch shapes are constants, include/kernel.h:9-10).  Times candidate
 * launch shapes on the given batch -- vote blocks per CU, finishing lanes,
 * one-sweep radix-sort tile -- and keeps the fastest on the context for
 * every later ppf_batch_lookup and server round.  Costs about three vote
 * passes and three finishing passes of the batch; results of every call are
 * identical whatever the shapes.  ppf_tune_reset restores the defaults. */
typedef struct {
  int32_t vote_blocks_per_cu; /* persistent vote blocks per compute unit */
  int32_t finish_lanes;       /* host threads + streams finishing jobs (0: all jobs together) */
  int32_t sort_tile_keys;     /* keys per one-sweep radix-sort tile */
  int32_t sort_tile_threads;  /* threads per one-sweep radix-sort tile */
  double best_vote_ms;        /* vote pass of the batch with the chosen shape */
  double best_finish_ms;      /* finishing of the batch with the chosen lanes */
  double best_sort_ms;        /* the tile's timed sorts */
  double tune_ms;             /* wall time of the call */
} ppf_tune_info_t;
int ppf_batch_tune(ppf_ctx ctx, int32_t n_jobs, const ppf_model *models, const ppf_scene *scenes,
                   const ppf_lookup_params *params, ppf_tune_info_t *info);
int ppf_tune_reset(ppf_ctx ctx);

/* ------------------------------------------------------------------ served
 * A request loop over resident models (no counterpart in the reference, whose
 * driver registers one pair at a time, src/cuda/ppf.cu:63-98).  A request is
 * one scene and a list of target model ids; the server copies it at submit,
 * admits it when its device-memory estimate fits the budget, uploads the
 * scene, and votes all admitted requests together in rounds (continuous
 * batching: requests arriving meanwhile join at the next round).  Finished
 * votes go through threshold/order, poses and clustering on the server's
 * lanes while the next rounds vote.  Each target's result is bit-identical to
 * ppf_lookup of the same (model, scene, params).  Models must outlive the
 * server.  All functions may be called from any thread. */
typedef struct ppf_server_s *ppf_server;
typedef struct {
  uint64_t hbm_budget_bytes; /* device memory for requests in flight; 0 = 80% of free memory */
  int32_t max_active;        /* requests in flight at once; 0 = 64 */
  int32_t lanes;             /* finishing lanes (host threads + streams); 0 = 4 */
  double round_ms;           /* target duration of one vote round; 0 = 50 */
} ppf_server_config;
typedef struct {
  int32_t status;      /* PPF_OK, or the first failing target's status */
  int32_t n_targets;
  int32_t vote_rounds; /* rounds in which this request was voted */
  int32_t pad;
  double queue_ms;     /* submit -> admission */
  double latency_ms;   /* submit -> all targets done */
  uint64_t device_bytes; /* admission estimate */
} ppf_request_info_t;
typedef struct {
  int64_t submitted, completed, failed, cancelled, rejected;
  int64_t rounds, rows_voted, queued, active;
  uint64_t inflight_bytes, peak_inflight_bytes, budget_bytes;
  int64_t max_requests_in_round; /* most distinct requests voted in one round */
  double vote_ms;
  int64_t finish_groups;   /* finishing calls of the lanes (a group: jobs finished together) */
  int64_t max_finish_group; /* most jobs finished together */
} ppf_server_stats_t;

int ppf_server_create(ppf_ctx ctx, int32_t n_models, const ppf_model *models,
                      const ppf_server_config *cfg, ppf_server *out);
/* Stops admitting, cancels queued requests, completes those in flight. */
int ppf_server_destroy(ppf_server server);
/* points/normals: float[n*3]; model_ids index the server's model list. */
int ppf_server_submit(ppf_server server, const float *points, const float *normals, int64_t n,
                      uint32_t ref_point_downsample_factor, int32_t n_targets,
                      const int32_t *model_ids, const ppf_lookup_params *params, uint64_t *ticket);
/* Wait up to timeout_s (< 0: no limit).  PPF_ERR_TIMEOUT leaves the request
 * pending; any other return consumes the ticket.  poses_out
 * [n_targets][16], status_out [n_targets], results_out [n_targets] (full
 * results, owned by the caller) and info may each be NULL.  Returns PPF_OK
 * when the request ran (per-target statuses in status_out), else its error. */
int ppf_server_wait(ppf_server server, uint64_t ticket, double timeout_s, float *poses_out,
                    int32_t *status_out, ppf_result *results_out, ppf_request_info_t *info);
int ppf_server_poll(ppf_server server, uint64_t ticket, int32_t *done);
/* Cancel a request that is still queued (PPF_ERR_STATE once admitted). */
int ppf_server_cancel(ppf_server server, uint64_t ticket);
int ppf_server_stats(ppf_server server, ppf_server_stats_t *stats);

/* Sizes of the ABI structs, in this order: ppf_model_info_t,
 * ppf_lookup_params, ppf_result_info_t, ppf_batch_info_t, ppf_server_config,
 * ppf_request_info_t, ppf_server_stats_t, ppf_tune_info_t.  Lets a binding check that it was
 * written against the same header. */
#define PPF_ABI_N_STRUCTS 8
void ppf_abi_sizes(uint32_t sizes_out[PPF_ABI_N_STRUCTS]);

/* ------------------------------------------------------------- diagnostics
 * Evaluate one numerics primitive on the device for n inputs (used to check
 * that device and host rounding agree).  op: 0 acos, 1 atan2(a,b), 2 sin,
 * 3 cos, 4 fmod(a,b), 5 quant_downf(a,b), 6 a/b, 7 sqrt(a), 8 alpha_bin as
 * float(a0,a1,b0,b1) from 4-vectors packed in a (n*2) and b (n*2). */
int ppf_debug_numerics(ppf_ctx ctx, int32_t op, const float *a, const float *b, float *out,
                       int64_t n);

/* Run one of the library's device primitives on host data (for tests):
 * op 0: exclusive scan of n uint32 (keys_io as uint32[n], result in place; the
 *       total is returned in vals_io[0]);
 * op 1: stable radix sort of n uint64 keys_io carrying uint32 vals_io (may be
 *       NULL) on bits [bit_lo, bit_hi);
 * op 2: plain device copy kernel of n uint32 (keys_io in, vals_io out), a
 *       known-size reference for hardware traffic counters;
 * op 3: n blocks of 256 threads each running bit_lo iterations of 16
 *       independent FMAs (keys_io receives n*256 floats), the full-rate VALU
 *       reference for instruction counters;
 * op 4: stable sort of n uint64 keys_io by their high 32 bits as the model
 *       build sorts its pair words: by the ranks of the distinct high halves
 *       when there are at most 8,192 of them and n >= 2^20 (vals_io[0] = 1),
 *       else by the radix sort (vals_io[0] = 0). */
int ppf_debug_primitive(ppf_ctx ctx, int32_t op, void *keys_io, uint32_t *vals_io, int64_t n,
                        int32_t bit_lo, int32_t bit_hi);

/* Diagnostic knobs of a context (for tests).  knob 0: the number of distinct
 * model buckets a reference row may group before the vote kernel votes that
 * row hit by hit (the path a row takes when its hits span more buckets than
 * the kernel's grouping table holds); 0 restores the table's size.  knob 1:
 * keys per thread of the one-sweep radix sort's tiles (8 or 16), knob 2:
 * threads per tile (256 or 512) (the autotuner's choice otherwise).  knob 3:
 * how ppf_batch_lookup finishes its jobs: 0 together when it can (default),
 * 1 job by job on lanes.  knob 4: ppf_batch_lookup's vote stage from captured
 * HIP graphs (1, default) or launched directly (0).  Results do not depend on
 * any of them. */
int ppf_debug_set(ppf_ctx ctx, int32_t knob, int64_t value);

/* GPU spatial-hash clustering (Model::ClusterTransformations + max_element,
 * src/cuda/model.cu:202-244) of host poses: T [P][16] row-major, codes [P]
 * (the model point m_r = bits 6..31 selects weights[m_r]; weights may be NULL
 * for all 1), counts [P].  Outputs vote_counts_out [P], transformation_trans
 * [P][3], transformation_rots [P][4]; stats_out = {max_idx, pair tests,
 * chunks spanning several cells that share a hash}. */
int ppf_debug_cluster(ppf_ctx ctx, const float *T, const uint64_t *codes, const uint32_t *counts,
                      const float *weights, int64_t n_weights, int64_t P, float d_dist, int32_t use_l1,
                      int32_t use_avg, float *vc_out, float *trans_out, float *quats_out,
                      int64_t stats_out[3]);

#ifdef __cplusplus
}
#endif
#endif /* PPF_AMD_H */
