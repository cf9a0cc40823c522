/*
 * mialign — MI355X (gfx950) seed -> extend -> sort path for linear WGBS reads.
 *
 * Plain C ABI: pointers and sizes only. Every call returns 0 on success or a
 * negative status; ma_last_error() then holds the message (the reference
 * raises a Mojo Error with the same text role). Host-pointer calls copy in and
 * out; *_dev calls take device pointers already resident in HBM and enqueue on
 * the given stream without synchronising.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   ma_index_open ........... LinearIndex._load_dense_pack + index upload,
 *                             fq2bam-meth/src/linear_index.mojo:217-303 and
 *                             linear_gpu_locate.mojo:1609-1703
 *   ma_index_build .......... build_dense_pack, fq2bam-meth/python/mojo_linear_pack.py:208-287
 *   ma_index_export_dense ... the dense-v1 files written by build_dense_pack (:271-273)
 *   ma_map_forward(_dev) .... the `-device cpu` host loop, linear_mapper.mojo:209-311
 *   ma_map_batch(_dev) ...... per-batch body of map_fastq_dense_gpu_locate,
 *                             linear_gpu_locate.mojo:1836-2720 (pack, FW/RC
 *                             encode+locate+vote_extend passes, FW/RC merge :2101-2113)
 *   ma_sort_keys_se_dev ..... SE key builders, linear_gpu_fm.mojo:371-405, 800-823
 *   ma_sort_keys_pe_dev ..... PE key builders, linear_gpu_fm.mojo:936-977
 *   ma_sort_markdup(_dev) ... gpu_sort_markdup, fq2bam-meth/src/linear_gpu_sort.mojo:24-41
 *   ma_seed_kmers ........... seed_kmers_on_device (pack + k-mer + Wang hash),
 *                             gpu-common/src/gpu_kernels.mojo:81-281
 *   ma_device_info .......... select_device / probe_device_context + hbm_info,
 *                             gpu-common/src/gpu_device.mojo:19-63, gpu_kernels.mojo:42-65,
 *                             gpu-common/python/gpu_mem.py:88-121
 *   ma_mate_rescue_dev ...... fm_mate_rescue_kernel, fq2bam-meth/src/linear_gpu_fm.mojo:1839-2011
 *   ma_bam_pack_se_dev ...... _pack_bam_aln over the SE branch of _emit_bam_from_batch
 *                             + _gather_perm_to_writer / _or_dup_flag,
 *                             fq2bam-meth/src/linear_gpu_fm.mojo:497-665, 755-800, 177-225, 408-420
 *   ma_bam_pack_pe_dev ...... _pack_bam_aln over the PE branch of _emit_bam_from_batch
 *                             (flags, mate fields, tlen), linear_gpu_fm.mojo:826-934
 *   ma_bam_pack_dev ......... _pack_bam_aln from explicit per-record fields (:497-665)
 *   ma_bgzf_* ............... _bgzf_block / _BgzfWriter / BamWriter.write_raw (host, zlib;
 *                             ma_bgzf_compress_dev: the same members compressed on the GPU),
 *                             fq2bam-meth/python/bam_emit.py:77-185, 477-481
 *   ma_fastq_parse_dev ...... the bulk FASTQ reader: _pipe_line / _store_rec / _qname_span /
 *                             _cvt / fq_pack_bases, fq2bam-meth/src/linear_fastq.mojo:37-46,
 *                             229-470 (FastqPairReader, fq2bam-meth/python/fastq_batch.py:230-317)
 *   ma_serve_* .............. no single reference counterpart: the request loop
 *                             generalises the batch loop of map_fastq_dense_gpu_locate
 *                             with admission by free HBM (gpu_mem.py:203-377)
 */
#ifndef MIALIGN_H
#define MIALIGN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MA_OK 0
#define MA_ERR_ARG (-1)
#define MA_ERR_HIP (-2)
#define MA_ERR_IO (-3)
#define MA_ERR_NOMEM (-4)
#define MA_ERR_UNSUPPORTED (-5)

typedef struct ma_index ma_index;
typedef struct ma_workspace ma_workspace;
typedef struct ma_server ma_server;

/* Mapping knobs; ma_params_default gives the reference defaults
 * (linear_gpu_locate.mojo:1729-1751, linear_index.mojo:425-434) except
 * max_indel, 0 (gapless extension). max_indel 1..4 (larger values clamp to 4)
 * turns on the single-indel branch (:1015-1024, 1142-1235, 1408-1511), the
 * reference's own default being 4: records then carry iop (1 insertion, 2
 * deletion), iat (query bases before the indel) and ilen. */
typedef struct {
    int32_t k;         /* k-mer length (index k) */
    int32_t stride;    /* seed stride (METHYLGRAPHER_LINEAR_SEED_STRIDE, 5) */
    int32_t max_occ;   /* locate cap (METHYLGRAPHER_LINEAR_MAX_OCC, 16384) */
    int32_t vote_occ;  /* vote / high-occ split (VOTE_OCC, 256) */
    int32_t vote_fast; /* pass-0 vote cap (VOTE_OCC_FAST, = vote_occ) */
    int32_t max_diff;  /* mismatch budget (MAX_DIFF, 6) */
    int32_t max_soft;  /* soft-clip cap (MAX_SOFT, 8) */
    int32_t max_indel; /* 0 = gapless; 1..4 = single-indel cap */
    int32_t do_rc;     /* reverse-complement pass (LINEAR_RC, 1) */
} ma_params;

/* One record per read, structure of arrays (vote_extend_kernel outputs,
 * linear_gpu_locate.mojo:778-787). cid -1 = unmapped; cid indexes the index's
 * contigs, so it also names the conversion copy (f* = C->T, r* = G->A). */
typedef struct {
    int32_t *cid;
    uint32_t *pos;  /* 0-based leftmost aligned base on the contig */
    uint32_t *mapq;
    int32_t *flag;  /* 0, 16 (reverse) or 4 (unmapped) */
    uint32_t *sl;   /* soft clip, left (reference orientation) */
    uint32_t *sr;   /* soft clip, right */
    uint32_t *nm;   /* mismatches */
    uint32_t *iop;  /* indel op: 0 none (always with max_indel 0), 1 insertion, 2 deletion */
    uint32_t *iat;  /* query bases before the indel */
    uint32_t *ilen; /* indel length (1..4) */
} ma_recs;

enum { MA_ENC_ASCII = 0, MA_ENC_2BIT = 1 };
enum { MA_BS_NONE = 0, MA_BS_C2T = 1, MA_BS_G2A = 2 };

/* ---- library / device ---- */
void ma_params_default(ma_params *p);
const char *ma_last_error(void);
const char *ma_version(void);
int ma_device_count(int *n);
int ma_device_info(int device, char *name, int name_len, int64_t *hbm_total, int64_t *hbm_free,
                   int32_t *n_cu, int32_t *lds_per_block);

/* Launch knobs, process-wide; every value gives bit-identical results (they
 * change tile sizes and launch shapes only). No reference counterpart: the
 * reference's batch knobs are hand-set (linear_gpu_locate.mojo:1729-1751);
 * here the autotuner (mialign.autotune) times candidates on the device and
 * sets these. ma_tune_get also returns the built-in default. */
enum {
    MA_TUNE_SORT_TILE_ITEMS = 0,     /* keys per thread of a radix tile: 8 or 16 (tiles of 2048 / 4096) */
    MA_TUNE_MAP_MAX_CLAIM = 1,       /* most reads a map wave claims per work-counter atomic, 1..64 */
    MA_TUNE_MAP_BLOCKS_PER_CU = 2,   /* persistent map grid: blocks (4 waves) per CU, 1..16 */
    MA_TUNE_MAP_CHUNKS_PER_WAVE = 3, /* default read chunks per resident wave of new workspaces, 1..64 */
    MA_TUNE_MAP_CLASS_SPLIT = 4,     /* 1: a batch of uneven reads maps each read-length class in its
                                        own launch (reads of <= 128, <= 160, <= 256 bases) */
    MA_TUNE_COUNT = 5
};
int ma_tune_set(int32_t knob, int64_t value);
int ma_tune_get(int32_t knob, int64_t *value, int64_t *default_value);

/* ---- device memory, streams and events (so hosts need no other GPU runtime) ---- */
int ma_dev_alloc(int device, int64_t bytes, void **ptr);
int ma_dev_free(void *ptr);
int ma_memcpy_h2d(void *dst, const void *src, int64_t bytes, void *stream);
int ma_memcpy_d2h(void *dst, const void *src, int64_t bytes, void *stream);
int ma_memcpy_d2d(void *dst, const void *src, int64_t bytes, void *stream);
int ma_stream_create(int device, void **stream);
int ma_stream_destroy(void *stream);
int ma_stream_sync(void *stream);
int ma_event_create(void **ev);
int ma_event_destroy(void *ev);
int ma_event_record(void *ev, void *stream);
int ma_event_sync(void *ev);
int ma_event_elapsed_ms(void *start, void *stop, float *ms);
int ma_event_query(void *ev); /* 1 = complete, 0 = pending, <0 error */
/* Stage chains as HIP graphs: ma_graph_begin starts capturing `stream` (this
 * thread's calls only); the library's *_dev calls enqueued on it until
 * ma_graph_end become the graph's nodes (kernels only: the device-pointer
 * calls make no allocation, copy, memset or synchronisation once their
 * buffers exist), which ma_graph_end instantiates (n_nodes may be NULL);
 * ma_graph_launch replays it on a stream. No reference counterpart (the
 * reference enqueues each kernel of a batch, linear_gpu_locate.mojo:1836-2753). */
int ma_graph_begin(void *stream);
int ma_graph_end(void *stream, void **graph_exec, int64_t *n_nodes);
int ma_graph_launch(void *graph_exec, void *stream);
int ma_graph_destroy(void *graph_exec);
/* Pinned host memory (page-locked; lets copies overlap kernels). */
int ma_host_alloc(int64_t bytes, void **ptr);
int ma_host_free(void *ptr);

/* ---- index ---- */
/* Load a dense-v1 directory (meta.json, kmers.bin, offsets.bin, postings.bin,
 * sequences.bin, contig_offsets.bin) and make it resident on `device`. */
int ma_index_open(const char *dense_v1_dir, int device, ma_index **out);
/* Same, from dense-v1 arrays already in host memory. */
int ma_index_from_dense(int device, int32_t k, const uint64_t *kmers, int64_t n_keys,
                        const uint64_t *offsets, const uint32_t *postings, int64_t n_postings,
                        const uint8_t *seq, int64_t n_bases, const uint64_t *contig_off,
                        int64_t n_contigs, ma_index **out);
/* Build the dense-v1 index on the GPU from (already converted) ASCII contig
 * bases; byte-identical to build_dense_pack's kmers/offsets/postings. The
 * windows are sorted one key range at a time, sized to the free HBM, so a
 * 3.1 Gbp x 2 reference builds within one GPU's memory. */
int ma_index_build(int device, int32_t k, const uint8_t *seq, int64_t n_bases,
                   const uint64_t *contig_off, int64_t n_contigs, ma_index **out);
typedef struct ma_build_opts {
    int32_t keep_dense;     /* 1 (default): kmers/offsets stay resident (export, table rebuild);
                               0: only the hash table, postings and reference (export_dense of
                               kmers/offsets then fails) */
    int32_t partition_bits; /* key-prefix partitions 2^bits the chunks are made of (1..10, default 10) */
    int64_t chunk_windows;  /* windows sorted at once (0 = from the free HBM) */
} ma_build_opts;
void ma_build_opts_default(ma_build_opts *opts);
int ma_index_build_opts(int device, int32_t k, const uint8_t *seq, int64_t n_bases,
                        const uint64_t *contig_off, int64_t n_contigs, const ma_build_opts *opts,
                        ma_index **out);
int ma_index_stats(const ma_index *ix, int64_t *n_keys, int64_t *n_postings, int64_t *n_bases,
                   int64_t *n_contigs, int64_t *device_bytes, int32_t *k);
/* Device hash table shape. slot_bytes 8: the compact table (64-byte buckets
 * of 8 slots, each a quotiented key tag + occ + start; n_overflow keys found
 * both of their buckets full and sit in a small stash); slot_bytes 16: 64-byte
 * buckets of 4 {key, occ << 40 | start} slots (n_overflow keys were placed by
 * the cuckoo pass). Every key sits in one of its two buckets (or the stash).
 * inline_postings 1: compact slots of one-posting keys hold that posting's
 * global base position instead of its start (a lookup needs no postings load;
 * indexes of fewer than 256 contigs whose positions fit the start field).
 * Any output pointer may be NULL. */
int ma_index_hash_stats(const ma_index *ix, int64_t *n_buckets, int64_t *n_overflow, int32_t *slot_bytes,
                        int32_t *inline_postings);
/* Copy the dense-v1 arrays back (kmers[n_keys], offsets[n_keys+1],
 * postings[2*n_postings]); any pointer may be NULL. */
int ma_index_export_dense(ma_index *ix, uint64_t *kmers, uint64_t *offsets, uint32_t *postings);
int ma_index_close(ma_index *ix);

/* ---- map ---- */
/* Workspace for up to max_reads reads of at most max_len bases. */
int ma_workspace_create(ma_index *ix, int64_t max_reads, int32_t max_len, ma_workspace **out);
int ma_workspace_destroy(ma_workspace *ws);
/* Launch policy of the map kernel for this workspace: 1 (default) = a
 * persistent grid (8 blocks per CU) taking read chunks from a counter; 0 = a
 * grid sized to the batch, one chunk per wave, so concurrent launches on other
 * streams interleave (what the served loop uses). Results are identical. */
int ma_workspace_set_persistent(ma_workspace *ws, int32_t persistent);
/* Launch shape: persistent as above; chunks_per_wave (1..64; until set, the
 * MA_TUNE_MAP_CHUNKS_PER_WAVE knob at each launch) = read chunks per resident
 * wave, so claim = clamp(n / (resident waves x chunks_per_wave), 1,
 * MA_TUNE_MAP_MAX_CLAIM) reads. More, smaller chunks shorten the tail of a
 * small batch. Results are identical for every setting. */
int ma_workspace_tune(ma_workspace *ws, int32_t persistent, int32_t chunks_per_wave);
/* Bytes a workspace of this shape holds on the device (for admission). */
int64_t ma_workspace_bytes(int64_t max_reads, int32_t max_len);

/* Reads in the reference's padded layout: read i at bases + i*stride, lens[i]
 * bases. encoding MA_ENC_ASCII (stride >= max_len, any byte; non-ACGT = N) or
 * MA_ENC_2BIT (4 bases/byte, base 4j in bits 7-6 of byte j, A0 C1 G2 T3).
 * In-silico conversion bs_r1 applies to reads [0, n_r1), bs_r2 to [n_r1, n)
 * (R1 then R2, linear_gpu_locate.mojo:1838-1876). Outputs: merged records
 * (required); fw / rc per-strand records (may be NULL). Host pointers. */
int ma_map_batch(ma_index *ix, const uint8_t *bases, int64_t stride, const uint32_t *lens,
                 int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                 int64_t n_r1, const ma_params *p, ma_recs *merged, ma_recs *fw, ma_recs *rc);
/* Device-pointer form; record arrays are device pointers; enqueued on stream. */
int ma_map_batch_dev(ma_index *ix, ma_workspace *ws, void *stream, const uint8_t *d_bases,
                     int64_t stride, const uint32_t *d_lens, int64_t n, int32_t max_len,
                     int32_t encoding, int32_t bs_r1, int32_t bs_r2, int64_t n_r1,
                     const ma_params *p, ma_recs *d_merged, ma_recs *d_fw, ma_recs *d_rc);

/* ma_map_batch_dev without a workspace, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc). No counterpart in the reference. Its one scratch buffer, the
 * persistent grid's work counter, is a static 256-byte device array of the
 * library (one per device, held until the process ends), so the call makes no
 * allocation of any kind; the counter is cleared by a kernel at each call.
 * Calls that overlap on one device share that counter and race: serialize
 * them (one stream, or events). Limits: n < 2^32 reads per call (the
 * counter's range), device ids 0..63; otherwise as ma_map_batch_dev with a
 * persistent workspace (same records). */
int ma_map_batch_dev_noalloc(ma_index *ix, void *stream, const uint8_t *d_bases, int64_t stride,
                             const uint32_t *d_lens, int64_t n, int32_t max_len, int32_t encoding,
                             int32_t bs_r1, int32_t bs_r2, int64_t n_r1, const ma_params *p,
                             ma_recs *d_merged, ma_recs *d_fw, ma_recs *d_rc);

/* The reference's `-device cpu` host loop on the device, for SE reads or one
 * mate at a time (fq2bam-meth/src/linear_mapper.mojo:209-311 ->
 * extend_read_with_seeds, linear_extend.mojo:143-211 -> vote_dense_seeds,
 * linear_index.mojo:530-578): forward strand only, no RC pass. ASCII rows
 * (stride >= max_len, max_len <= 256), converted as that loop does (for C2T a
 * C or c becomes T, for G2A a G or g becomes A; bs_r1 for reads [0, n_r1),
 * bs_r2 after). Every k-mer position i of a read (no byte outside ACGT/acgt,
 * in the index, at most max_occ postings) votes for (contig, pos - i *
 * p->stride) with each posting at pos >= i * p->stride; the first key whose
 * count exceeds every count before it wins; its window of the read's length
 * (none past the contig's end) is compared byte for byte (upper case; a
 * non-ACGT reference base reads as N); over p->max_diff mismatches the
 * leading and trailing mismatch runs (p->max_soft each at most) are clipped
 * and the rest recounted. Mapped when at most p->max_diff mismatches remain
 * over at least 32 aligned bases: pos = start + left clip, flag 0, MAPQ 60 /
 * 40 / 20 for >= 5 / >= 3 / fewer votes, nm the count; iop / iat / ilen 0.
 * Other p fields are not used; p->k must be the index's. Host pointers. */
int ma_map_forward(ma_index *ix, const uint8_t *bases, int64_t stride, const uint32_t *lens, int64_t n,
                   int32_t max_len, int32_t bs_r1, int32_t bs_r2, int64_t n_r1, const ma_params *p,
                   ma_recs *out);
/* ma_map_forward on device pointers. Two launches on stream with a host wait
 * between them: the first finds the batch's largest count of voting postings
 * of one read, which sizes the vote tables (one per wave, at most 4 GiB
 * together, grown in ws and kept). Not for graph capture. */
int ma_map_forward_dev(ma_index *ix, ma_workspace *ws, void *stream, const uint8_t *d_bases, int64_t stride,
                       const uint32_t *d_lens, int64_t n, int32_t max_len, int32_t bs_r1, int32_t bs_r2,
                       int64_t n_r1, const ma_params *p, ma_recs *d_out);

/* Mate rescue of a paired batch, in place on its merged records (device
 * pointers; R1 block [0, n_pairs) with bs_r1, R2 block [n_pairs, 2 n_pairs)
 * with bs_r2, rows as ma_map_batch_dev's). For a pair with exactly one mate
 * mapped, the other is scanned gapless within +-window (clamped to 32..1024)
 * of the mate on the mate's contig, the opposite strand first; the first
 * least mismatch count within clamp(max_diff, 2, len / 2) wins (counting stops
 * past budget + 4); with none, a one-base deletion at the first least window,
 * split points every 4 bases. A rescued mate: the contig, the window start,
 * flag 16 if reverse, no clips, nm, MAPQ 60 / 40 / 20 (nm 0 / <= 2 / more),
 * iop / iat / ilen 0; reads shorter than 24 stay unmapped; window <= 0 does
 * nothing. Replaces fm_mate_rescue_kernel of the native-BAM driver,
 * fq2bam-meth/src/linear_gpu_fm.mojo:1839-2011 (METHYLGRAPHER_FM_RESCUE_WIN,
 * default 512, :2082). Enqueued on stream. */
int ma_mate_rescue_dev(ma_index *ix, void *stream, const uint8_t *d_bases, int64_t stride, const uint32_t *d_lens,
                       int64_t n_pairs, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                       int32_t window, int32_t max_diff, ma_recs *d_merged);

/* ---- sort keys, sort, duplicate marking ---- */
/* SE keys from merged records (device pointers): coord, dhi, dlo, score, pair
 * (pair = pair_base + i). d_tid_of_cid[n_tid] maps index contig -> canonical
 * target; index contigs outside [0, n_tid) are unmapped. */
int ma_sort_keys_se_dev(void *stream, int64_t n, const ma_recs *d_merged, const uint32_t *d_lens,
                        const int32_t *d_tid_of_cid, int32_t n_tid, int64_t pair_base, uint64_t *d_coord,
                        uint64_t *d_dhi, uint64_t *d_dlo, uint32_t *d_score, uint32_t *d_pair);
/* PE keys of n_pairs pairs (device pointers): the merged records and lens hold
 * R1 block then R2 block (2 n_pairs each); the 2 n_pairs keys come out in
 * emission order, pair p's R1 at 2p and R2 at 2p + 1. coord per record; dhi /
 * dlo (the pair's unclipped 5' ends, ordered), score ((mapq1 + mapq2) << 8 |
 * 255 - min(nm1 + nm2, 255)) and pair ((pair_base + 2p) / 2) shared by the pair.
 * Index contigs outside [0, n_tid) are unmapped. */
int ma_sort_keys_pe_dev(void *stream, int64_t n_pairs, const ma_recs *d_merged, const uint32_t *d_lens,
                        const int32_t *d_tid_of_cid, int32_t n_tid, int64_t pair_base, uint64_t *d_coord,
                        uint64_t *d_dhi, uint64_t *d_dlo, uint32_t *d_score, uint32_t *d_pair);
/* gpu_sort_markdup contract: is_dup (original order), perm (stable coord
 * order), dup_perm (stable (dhi, dlo) order, identity when !do_markdup; may be
 * NULL). Host pointers. */
int ma_sort_markdup(int device, int64_t n, const uint64_t *coord, const uint64_t *dhi,
                    const uint64_t *dlo, const uint32_t *score, const uint32_t *pair,
                    uint32_t *is_dup, uint32_t *perm, int32_t do_markdup, uint32_t *dup_perm);
/* Device form; scratch sized by ma_sort_scratch_bytes(n) (caller-owned). */
int64_t ma_sort_scratch_bytes(int64_t n);
int ma_sort_markdup_dev(void *stream, int64_t n, const uint64_t *d_coord, const uint64_t *d_dhi,
                        const uint64_t *d_dlo, const uint32_t *d_score, const uint32_t *d_pair,
                        uint32_t *d_is_dup, uint32_t *d_perm, int32_t do_markdup,
                        uint32_t *d_dup_perm, void *d_scratch, int64_t scratch_bytes);
/* After ma_sort_markdup_dev on `stream` (same n and scratch): synchronizes the
 * stream and reports MA_ERR_HIP if a radix pass gave up waiting on a preceding
 * tile's published counts (a bounded wait that never occurs in a correct run;
 * its results would then be invalid), MA_OK otherwise. */
int ma_sort_dev_status(void *stream, int64_t n, const void *d_scratch, int64_t scratch_bytes);
/* Stable LSD radix sort of u64 keys carrying u32 values (device pointers). */
int ma_radix_sort_u64_dev(void *stream, int64_t n, uint64_t *d_keys, uint32_t *d_vals,
                          void *d_scratch, int64_t scratch_bytes);

/* ---- BAM records and BGZF ---- */
/* Per-record BAM fields (device int32 arrays); ntid / npos / tlen may be NULL
 * (mate tid -1, mate pos -1, tlen 0). Meaning as _pack_bam_aln's arguments. */
typedef struct {
    const int32_t *flag, *tid, *pos, *mapq, *sl, *sr, *nm, *ntid, *npos, *tlen;
} ma_bam_fields;
/* The reads of a batch (device pointers): read i's name is
 * names[name_off[i] .. name_off[i+1]) (no terminator), its original
 * (pre-conversion) ASCII bases seq + i*stride (lens[i] of them) and its FASTQ
 * quality qual + i*stride (qual_lens[i] bytes, qual_lens NULL = lens; qual
 * NULL = absent, QUAL 0xFF). */
typedef struct {
    const uint8_t *names;
    const int64_t *name_off;
    const uint8_t *seq;
    const uint8_t *qual;
    const uint32_t *lens;
    const uint32_t *qual_lens;
    int64_t stride;
} ma_bam_reads;
/* Device scratch for n records; bytes one record takes at most. */
int64_t ma_bam_scratch_bytes(int64_t n);
int64_t ma_bam_record_bound(int32_t name_len, int32_t len, int32_t rg_n);
/* Uncompressed BAM alignment records (block_size first, little endian) of n
 * single-end merged map records, written back to back into d_out in output
 * order: output record j is read perm[j] (d_perm NULL = identity; otherwise a
 * permutation of [0, n) — other values give empty records, never a fault), with flag
 * 0x400 set when apply_dup and d_is_dup[perm[j]] != 0. Fields as the
 * reference's SE batch emit: flag = record flag | 0x41, tid = d_tid_of_cid[cid]
 * (cid outside [0, n_tid) -> unmapped, flag | 4), CIGAR slS mM srS (clips
 * swapped for flag 16), SEQ / QUAL reversed (and complemented) for flag 16,
 * tags RG:Z:<rg> NM:C. d_rec_off[n+1] receives each record's byte offset and
 * the total in d_rec_off[n]; records that would pass out_cap are not written
 * (the caller compares d_rec_off[n] with out_cap). d_out 4-byte aligned;
 * rg_n <= 240. Enqueued on stream. */
int ma_bam_pack_se_dev(void *stream, int64_t n, const ma_recs *d_merged, const int32_t *d_tid_of_cid,
                       int32_t n_tid, const ma_bam_reads *reads, const char *rg, int32_t rg_n,
                       const uint32_t *d_perm, const uint32_t *d_is_dup, int32_t apply_dup,
                       uint8_t *d_out, int64_t out_cap, uint64_t *d_rec_off, void *d_scratch,
                       int64_t scratch_bytes);
/* Paired-end records of n_pairs pairs: merged records and reads R1 block then
 * R2 block (2 n_pairs each); the 2 n_pairs records in emission order (pair p's
 * R1 at 2p, R2 at 2p + 1) are what d_perm / d_is_dup index. Fields as the
 * reference's PE batch emit: flag | 0x41 (R1) or | 0x81 (R2), 0x8 when the mate
 * has no target, both on one target: 0x2, mate tid / pos and tlen = |pos2 -
 * pos1| + len(R2), positive on the leftmost mate; else a mapped mate's tid /
 * pos in the other's mate fields. Otherwise as ma_bam_pack_se_dev. */
int ma_bam_pack_pe_dev(void *stream, int64_t n_pairs, const ma_recs *d_merged, const int32_t *d_tid_of_cid,
                       int32_t n_tid, const ma_bam_reads *reads, const char *rg, int32_t rg_n,
                       const uint32_t *d_perm, const uint32_t *d_is_dup, int32_t apply_dup,
                       uint8_t *d_out, int64_t out_cap, uint64_t *d_rec_off, void *d_scratch,
                       int64_t scratch_bytes);
/* The same from explicit fields (the generic _pack_bam_aln batch). */
int ma_bam_pack_dev(void *stream, int64_t n, const ma_bam_fields *fields, const ma_bam_reads *reads,
                    const char *rg, int32_t rg_n, const uint32_t *d_perm, const uint32_t *d_is_dup,
                    int32_t apply_dup, uint8_t *d_out, int64_t out_cap, uint64_t *d_rec_off,
                    void *d_scratch, int64_t scratch_bytes);
/* BGZF: consecutive 65280-byte members (raw deflate at `level`, zlib) of a
 * host stream, compressed by `threads` workers, optionally followed by the
 * 28-byte EOF member; ma_bgzf_bound(n) bytes always suffice. */
int64_t ma_bgzf_bound(int64_t n);
int ma_bgzf_compress(const uint8_t *src, int64_t n, int32_t level, int32_t threads, int32_t add_eof,
                     uint8_t *dst, int64_t dst_cap, int64_t *out_n);
/* BGZF inflate (replaces the reference's gzip FASTQ pipe, `pigz -dc` or
 * `gzip -dc`, fq2bam-meth/src/linear_fastq.mojo:141-150, for bgzip-compressed
 * files): *used = the bytes of the whole members at the front of src[0..n)
 * (each a gzip header whose extra field holds BC = member bytes - 1, raw
 * deflate, CRC32, ISIZE), *out_n = their inflated size. With dst null only
 * those two are set; else the members are inflated into dst (dst_cap >=
 * *out_n) by `threads` workers, each member's ISIZE and CRC32 checked.
 * MA_ERR_ARG for bytes that are not a BGZF member, MA_ERR_IO for a corrupt one. */
int ma_bgzf_inflate(const uint8_t *src, int64_t n, int32_t threads, uint8_t *dst, int64_t dst_cap,
                    int64_t *used, int64_t *out_n);
/* gzip FASTQ input of any framing (replaces the same `pigz -dc` / `gzip -dc`
 * pipe, fq2bam-meth/src/linear_fastq.mojo:141-150): zlib's gz reader over a
 * file, every member inflated in turn, each CRC32 and ISIZE checked.
 * ma_gz_read fills dst with up to cap inflated bytes (*got = 0 at the end);
 * MA_ERR_ARG if the file is not gzip, MA_ERR_IO for a corrupt or truncated one. */
typedef struct ma_gz ma_gz;
int ma_gz_open(const char *path, ma_gz **out);
int ma_gz_read(ma_gz *f, uint8_t *dst, int64_t cap, int64_t *got);
int ma_gz_close(ma_gz *f);
/* BGZF on the device: the same 65280-byte members (header, raw deflate body,
 * CRC32, ISIZE) of a device stream, compressed by the GPU: level 0 = stored
 * blocks, otherwise LZ77 (greedy, matches of 4..64 bytes) and one
 * dynamic-Huffman block per member (a stored block where that is not smaller). The bodies are this library's own deflate, so the
 * bytes differ from zlib's; the members decode to the same stream. Writes
 * d_dst (ma_bgzf_bound(n) bytes suffice; 4-byte aligned, like d_src) and the
 * total length to d_total (device); members past dst_cap are not written.
 * Scratch: ma_bgzf_dev_scratch_bytes(n). Enqueued on stream. */
int64_t ma_bgzf_dev_scratch_bytes(int64_t n);
int ma_bgzf_compress_dev(void *stream, const uint8_t *d_src, int64_t n, int32_t level, int32_t add_eof,
                         uint8_t *d_dst, int64_t dst_cap, uint64_t *d_total, void *d_scratch,
                         int64_t scratch_bytes);
/* A BGZF file writer: writes accumulate into 65280-byte members exactly as one
 * stream (a member never waits on a record boundary); close flushes the last
 * partial member and the EOF member. */
typedef struct ma_bgzf ma_bgzf;
int ma_bgzf_open(const char *path, int32_t level, int32_t threads, ma_bgzf **out);
int ma_bgzf_write(ma_bgzf *w, const void *data, int64_t n);
int ma_bgzf_close(ma_bgzf *w, int64_t *bytes_written);

/* ---- FASTQ ingest ---- */
/* Parses raw FASTQ bytes resident in HBM into a read batch, as the reference's
 * bulk reader does (lines end at '\n', one trailing '\r' dropped; a record is
 * four lines; the name skips '@' and stops at ' ' or '/'; bs_mode 1 = C2T,
 * C/c -> T, 2 = G2A, G/g -> A, 0 = none). Up to max_records records: d_rows
 * [max_records x stride] original bases padded with 'N'; d_conv_rows the
 * converted bases (may be NULL); d_qual_rows the first `stride` quality bytes,
 * 0 padded (may be NULL); d_lens / d_qual_lens; names back to back in d_names
 * at d_name_off[max_records + 1]. d_info[8] (device u64): records, bytes
 * consumed (through the last record's quality line), longest read, status
 * (0 ok; 1 / 2 / 3 = the input ends inside a record, missing SEQ / '+' / QUAL,
 * only when `final`; 4 = a read longer than stride; 5 = names past names_cap —
 * the batch stops before that record), name bytes, newlines, lines. Without
 * `final` an unterminated or incomplete tail stays unconsumed for the next
 * buffer. Scratch: ma_fastq_scratch_bytes. Enqueued on stream. */
int64_t ma_fastq_scratch_bytes(int64_t nbytes, int64_t max_records);
int ma_fastq_parse_dev(void *stream, const uint8_t *d_buf, int64_t nbytes, int32_t final, int64_t max_records,
                       int32_t stride, int32_t bs_mode, uint8_t *d_rows, uint8_t *d_conv_rows, uint8_t *d_qual_rows,
                       uint32_t *d_lens, uint32_t *d_qual_lens, uint8_t *d_names, int64_t names_cap,
                       int64_t *d_name_off, uint64_t *d_info, void *d_scratch, int64_t scratch_bytes);

/* ---- gpu-common seed kernels ---- */
/* ASCII reads [n x max_len] -> per-position Wang hash of the MSB-first k-mer
 * (0 = invalid window), out_hash[n*max_len]. Host pointers. */
int ma_seed_kmers(int device, const uint8_t *bases, int64_t n, int32_t max_len, int32_t k,
                  uint64_t *out_hash);

/* ---- served request loop ---- */
/* A server owns `slots` streams and maps queued read batches (requests) with
 * continuous batching: requests are cut into pieces of at most `launch_reads`
 * reads (default 131072; a paired request is also split where its conversion
 * changes) and, whenever a slot is free, one launch is assembled from waiting
 * pieces in deadline order (see ma_serve_tune), merging compatible pieces (same
 * row layout, conversion and knobs).
 * Device memory: everything the server holds (slot buffers, map workspaces,
 * the target table, sorted requests' arenas) stays within hbm_budget bytes
 * (0 = no cap of its own), and every time it takes more it first reads the
 * device's free HBM (hipMemGetInfo) and takes it only if 256 MiB stay free.
 * Slot buffers are reserved here; steady-state launches allocate nothing.
 * Replaces no single reference call: the admission follows
 * gpu-common/python/gpu_mem.py:203-377 (free-HBM preflight and wait). */
int ma_serve_open(ma_index *ix, int32_t slots, int64_t hbm_budget, int32_t max_len,
                  ma_server **out);
/* launch_reads >= 1024 caps the reads of one launch; waiting work runs in
 * order of arrival + size_weight x its estimated stand-alone time (0 = arrival
 * order; default 0.5); persistent / chunks_per_wave set every slot's launch
 * shape as in ma_workspace_tune (defaults 0 and 4). Only while nothing is in
 * flight. */
int ma_serve_tune(ma_server *s, int64_t launch_reads, double size_weight, int32_t persistent,
                  int32_t chunks_per_wave);
/* Target table for sorted requests (host int32 tid_of_cid[n_tid], as
 * ma_sort_keys_se_dev takes it; default: index contig c is target c). Only
 * while no request is queued or running. */
int ma_serve_set_targets(ma_server *s, const int32_t *tid_of_cid, int32_t n_tid);
/* Queue host reads (same layout and record outputs as ma_map_batch's merged
 * records); the host arrays must stay valid until the ticket completes (pinned
 * memory lets the copies overlap; ten record arrays at one constant pitch in
 * one block are copied back with one 2D copy per piece). Returns a ticket >= 0,
 * or MA_ERR_NOMEM when the request is not admitted now (nothing of it queued). */
int64_t ma_serve_submit(ma_server *s, const uint8_t *bases, int64_t stride, const uint32_t *lens,
                        int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                        int64_t n_r1, const ma_params *p, ma_recs *merged_out);
/* As ma_serve_submit, and the request is one sample's batch that is sorted and
 * duplicate-marked on its own once mapped (the reference marks duplicates per
 * sample: fq2bam-meth/src/linear_gpu_fm.mojo:800-823 builds the keys of one
 * sample's records, linear_gpu_sort.mojo:156-551 sorts and marks them).
 * keys 1: SE keys over its n records; keys 2: PE keys, R1 rows [0, n/2) and R2
 * rows [n/2, n) with n_r1 = n/2, records then numbered in emission order (pair
 * p's R1 = 2p, R2 = 2p + 1, as ma_sort_keys_pe_dev). is_dup[n] and perm[n]
 * (host, valid until completion) receive what ma_sort_markdup_dev gives for
 * those keys (pair_base 0). Its device arena (~112 bytes per read) is taken at
 * submit: MA_ERR_NOMEM when it does not fit now (wait for completions and
 * submit again). A sort whose radix pass gave up waiting is reported as
 * MA_ERR_HIP by the poll / reap call that completes the ticket. */
int64_t ma_serve_submit_sorted(ma_server *s, const uint8_t *bases, int64_t stride, const uint32_t *lens,
                               int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                               int64_t n_r1, const ma_params *p, ma_recs *merged_out, int32_t keys,
                               int32_t do_markdup, uint32_t *is_dup, uint32_t *perm);
/* Device bytes the server holds, its budget (0 = none) and the device's free
 * HBM now. Any pointer may be NULL. */
int ma_serve_memory(ma_server *s, int64_t *held_bytes, int64_t *budget, int64_t *free_hbm);
/* Returns 1 when the ticket is complete (records written to its arrays) and
 * writes its latency (submit to observed completion, ms), 0 while it is queued
 * or running; wait != 0 blocks until complete. */
int ma_serve_poll(ma_server *s, int64_t ticket, int32_t wait, double *latency_ms);
/* Progress the server and report finished tickets: writes up to `cap` of them
 * to `tickets` and returns how many (each is reported once, by this call or by
 * ma_serve_poll). wait != 0 blocks until one finishes or nothing is in flight.
 * MA_ERR_NOMEM when nothing is in flight and the most urgent queued request
 * cannot get a launch buffer. */
int ma_serve_reap(ma_server *s, int64_t *tickets, int32_t cap, int32_t wait);
int ma_serve_close(ma_server *s);

#ifdef __cplusplus
}
#endif
#endif
