/*
 * ORACLE — CPU restatement of the reference's linear WGBS seed -> extend ->
 * sort path. TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library as the checker. The product
 * (mojo-align_amd/) never links, loads or calls it.
 *
 * The reference is Mojo (fq2bam-meth/src/ .mojo files) and cannot be compiled here,
 * so every function below restates the reference's arithmetic from its source
 * text; each cites the file:line it follows. Knob defaults are the reference's
 * (linear_gpu_locate.mojo:1729-1751): k=21 (caller), stride 5, max_occ 16384,
 * vote_occ = vote_fast = 256, max_diff 6, max_soft 8, RC on; indel_cap is a
 * parameter (0 = gapless contract).
 *
 * Parity status per function (how each is pinned by tests/test_oracle.py):
 *   or_code_of / or_pack_bases / or_bs_convert ..... reference known answers +
 *                                                    closed form (per-base table)
 *   or_wang_hash ..................................... published integer hash;
 *                                                    pinned by bijectivity /
 *                                                    avalanche checks only
 *   or_encode_strided ................................ brute-force string k-mers
 *   or_build_dense / or_locate_bsearch ............... brute-force window scan
 *   or_vote_extend / or_map_batch .................... PARITY UNPINNED by any
 *        reference vector (none exists); pinned by invariants, brute-force
 *        best-locus search on unique references and a hand-computed branch suite
 *   or_sort_keys_se .................................. closed-form key formulas
 *   or_sort_markdup .................................. reference known answers
 *                                                    (test_bam_emit.py) + brute force
 *   or_reg2bin / or_nt16 / or_nt16_comp ............. reference known answers
 *        (test_bam_emit.py) + the SAM specification's bin definition by brute force
 *   or_pack_bam_aln / or_bam_pack_fields / or_bam_pack_se ... reference known
 *        answers for the 4-bit SEQ and the bin + an independent BAM decoder
 *        (tests/bam_decode.py) checked field by field against the SAM text of
 *        the same records + planted mistakes (swapped clips, unreversed QUAL)
 *   or_pe_fields / or_bam_pack_pe ..................... hand-worked pairs (one
 *        contig in both orders, two contigs, one or both mates unmapped), the
 *        parity path's own pairing (pair_hits, linear_extend.mojo:263-294) on
 *        flags, mate fields and |tlen|, and decoded records whose mate fields
 *        name their mate's tid / pos
 *   or_fq_cvt / or_fastq_parse ........................ reference known answers
 *        (test_linear_fastq.mojo, test_fastq_batch.py, the reference's FASTQ
 *        fixture) + the host reader of the parity path (read_one) on random
 *        FASTQ text + hand-cut edge cases (CRLF, missing final newline, truncated
 *        records, empty lines)
 *   or_recount_nm (a checker, no reference counterpart) ... an independent Python
 *                                                    recount + planted one-base shifts
 *   or_mate_rescue .................................... hand-derived pairs (test_rescue.py)
 *   or_map_forward (the `-device cpu` host loop) ....... hand-derived reads (unique,
 *        mismatches, broken first k-mer, clipping, short reads, contig end, N, case
 *        and conversion) + the vote written as the earliest max-count-th occurrence
 *        on repeat reads over all MAPQ bands (test_forward.py)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_EMPTY_KEY 0xFFFFFFFFFFFFFFFFull

/* ------------------------------------------------------------------ A1/A2 */

/* ASCII -> 2-bit code; anything but A/C/G/T (either case) -> 255.
 * Follows pack_bases_kernel, gpu_kernels.mojo:108-126 (= linear_gpu_locate.mojo:399-417). */
uint8_t or_code_of(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 255;
    }
}

void or_pack_bases(const uint8_t *bases, uint8_t *codes, int64_t n) {
    for (int64_t i = 0; i < n; i++) codes[i] = or_code_of(bases[i]);
}

/* In-silico bisulfite conversion of ASCII text. mode 1 = C2T (Cc -> Tt),
 * mode 2 = G2A (Gg -> Aa); other bytes untouched.
 * Follows _bs_convert, linear_gpu_locate.mojo:58-70 (str.translate tables). */
void or_bs_convert(uint8_t *seq, int64_t n, int mode) {
    for (int64_t i = 0; i < n; i++) {
        uint8_t b = seq[i];
        if (mode == 1) {
            if (b == 'C') seq[i] = 'T';
            else if (b == 'c') seq[i] = 't';
        } else if (mode == 2) {
            if (b == 'G') seq[i] = 'A';
            else if (b == 'g') seq[i] = 'a';
        }
    }
}

/* Reference-FASTA conversion: only C/c -> T (uppercase T even for 'c').
 * Follows convert_fasta_c2t, fq2bam_meth.py:79-88. */
void or_convert_fasta_c2t(uint8_t *seq, int64_t n) {
    for (int64_t i = 0; i < n; i++)
        if (seq[i] == 'C' || seq[i] == 'c') seq[i] = 'T';
}

/* ------------------------------------------------------------------ A3 */

/* Thomas Wang 64-bit integer mix; follows wang_hash_u64, gpu_kernels.mojo:128-137. */
uint64_t or_wang_hash(uint64_t key) {
    key = (~key) + (key << 21);
    key = key ^ (key >> 24);
    key = (key + (key << 3)) + (key << 8);
    key = key ^ (key >> 14);
    key = (key + (key << 2)) + (key << 4);
    key = key ^ (key >> 28);
    key = key + (key << 31);
    return key;
}

/* Strided k-mer keys, MSB-first 2-bit; ~0 when the window overruns the read or
 * holds a code > 3. slot s -> q_off = s*stride.
 * Follows encode_strided_kernel, linear_gpu_locate.mojo:419-451. */
static void encode_one_read(const uint8_t *codes, int64_t rid, uint32_t L, int64_t max_len,
                            int64_t k, int64_t stride, int64_t slots_per,
                            uint64_t *keys, uint32_t *qoffs) {
    for (int64_t s = 0; s < slots_per; s++) {
        int64_t idx = rid * slots_per + s;
        int64_t q_off = s * stride;
        keys[idx] = OR_EMPTY_KEY;
        qoffs[idx] = (uint32_t)q_off;
        if (q_off + k > (int64_t)L) continue;
        const uint8_t *p = codes + rid * max_len + q_off;
        uint64_t key = 0;
        int ok = 1;
        for (int64_t j = 0; j < k; j++) {
            if (p[j] > 3) { ok = 0; break; }
            key = (key << 2) | p[j];
        }
        if (ok) keys[idx] = key;
    }
}

void or_encode_strided(const uint8_t *codes, const uint32_t *lens, int64_t n_reads,
                       int64_t max_len, int64_t k, int64_t stride, int64_t slots_per,
                       uint64_t *keys, uint32_t *qoffs) {
    for (int64_t r = 0; r < n_reads; r++)
        encode_one_read(codes, r, lens[r], max_len, k, stride, slots_per, keys, qoffs);
}

/* ------------------------------------------------------------------ A4 */

/* lower_bound over sorted unique kmers -> CSR [start, end); zeroed when the key
 * is absent, invalid, or occ > max_occ.
 * Follows locate_bsearch_kernel, linear_gpu_locate.mojo:453-490 and
 * _dense_find_key, linear_index.mojo:436-453. */
static void locate_one(uint64_t key, const uint64_t *kmers, const uint64_t *offsets,
                       int64_t n_table, int64_t max_occ, uint64_t *st, uint64_t *en) {
    *st = 0;
    *en = 0;
    if (key == OR_EMPTY_KEY || n_table <= 0) return;
    int64_t lo = 0, hi = n_table;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (kmers[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= n_table || kmers[lo] != key) return;
    uint64_t a = offsets[lo], b = offsets[lo + 1];
    int64_t occ = (int64_t)(b - a);
    if (occ <= 0 || occ > max_occ) return;
    *st = a;
    *en = b;
}

void or_locate_bsearch(const uint64_t *keys, int64_t n_slots, const uint64_t *kmers,
                       const uint64_t *offsets, int64_t n_table, int64_t max_occ,
                       uint64_t *out_start, uint64_t *out_end) {
    for (int64_t i = 0; i < n_slots; i++)
        locate_one(keys[i], kmers, offsets, n_table, max_occ, &out_start[i], &out_end[i]);
}

/* ------------------------------------------------- dense-v1 index build */

/* Number of k-mer windows (before N filtering) over all contigs. */
int64_t or_count_windows(const uint64_t *contig_off, int64_t n_contigs, int k) {
    int64_t n = 0;
    for (int64_t c = 0; c < n_contigs; c++) {
        int64_t len = (int64_t)(contig_off[c + 1] - contig_off[c]);
        if (len >= k) n += len - k + 1;
    }
    return n;
}

typedef struct { uint64_t key; uint64_t loc; } or_kl;

/* Stable merge sort on key (equal keys keep input order). */
static void merge_sort_kl(or_kl *a, or_kl *tmp, int64_t n) {
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n;
            int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) tmp[o++] = (a[j].key < a[i].key) ? a[j++] : a[i++];
            while (i < mid) tmp[o++] = a[i++];
            while (j < hi) tmp[o++] = a[j++];
        }
        memcpy(a, tmp, (size_t)n * sizeof(or_kl));
    }
}

/* dense-v1 build over ASCII contig bases (already converted; case-folded to
 * upper by the reader). kmers: sorted unique MSB-first 2-bit keys; offsets:
 * CSR (n_keys+1); postings: u32 pairs (contig_id, pos0) ordered by (key, input
 * order) = (key, cid, pos) because contigs are enumerated in order.
 * Follows encode_kmers_contig + build_dense_pack, mojo_linear_pack.py:179-287
 * (stable mergesort :239). Capacities: kmers/offsets-1/postings/2 >=
 * or_count_windows(). Returns 0 or -1 on allocation failure. */
int or_build_dense(const uint8_t *seq, const uint64_t *contig_off, int64_t n_contigs, int k,
                   uint64_t *kmers_out, uint64_t *offsets_out, uint32_t *postings_out,
                   int64_t *n_keys_out, int64_t *n_postings_out) {
    int64_t cap = or_count_windows(contig_off, n_contigs, k);
    or_kl *a = (or_kl *)malloc((size_t)(cap > 0 ? cap : 1) * sizeof(or_kl));
    or_kl *t = (or_kl *)malloc((size_t)(cap > 0 ? cap : 1) * sizeof(or_kl));
    if (!a || !t) { free(a); free(t); return -1; }
    uint64_t kmask = (k >= 32) ? ~0ull : ((1ull << (2 * k)) - 1ull);
    int64_t n = 0;
    for (int64_t c = 0; c < n_contigs; c++) {
        int64_t off = (int64_t)contig_off[c];
        int64_t len = (int64_t)(contig_off[c + 1] - contig_off[c]);
        if (len < k) continue;
        for (int64_t p = 0; p + k <= len; p++) {
            uint64_t key = 0;
            int ok = 1;
            for (int j = 0; j < k; j++) {
                uint8_t cd = or_code_of(seq[off + p + j]);
                if (cd > 3) { ok = 0; break; }
                key = (key << 2) | cd;
            }
            if (!ok || key > kmask) continue;
            a[n].key = key;
            a[n].loc = ((uint64_t)c << 32) | (uint64_t)p;
            n++;
        }
    }
    merge_sort_kl(a, t, n);
    int64_t nk = 0;
    for (int64_t i = 0; i < n; i++) {
        if (i == 0 || a[i].key != a[i - 1].key) {
            kmers_out[nk] = a[i].key;
            offsets_out[nk] = (uint64_t)i;
            nk++;
        }
        postings_out[2 * i] = (uint32_t)(a[i].loc >> 32);
        postings_out[2 * i + 1] = (uint32_t)(a[i].loc & 0xFFFFFFFFull);
    }
    offsets_out[nk] = (uint64_t)n;
    *n_keys_out = nk;
    *n_postings_out = n;
    free(a);
    free(t);
    return 0;
}

/* ------------------------------------------------------------ A6 / A7 */

typedef struct {
    const uint64_t *kmers;
    const uint64_t *offsets;
    const uint32_t *postings; /* (cid, pos) pairs */
    int64_t n_keys;
    int64_t n_postings;
    const uint8_t *ref_codes; /* pack_bases(sequences.bin) */
    const uint64_t *contig_off;
    int64_t n_contigs;
} or_index;

typedef struct {
    int32_t k;
    int32_t stride;
    int32_t max_occ;
    int32_t vote_occ;
    int32_t vote_fast;
    int32_t max_diff;
    int32_t max_soft;
    int32_t max_indel;
    int32_t do_rc;
} or_params;

/* Output record arrays (SoA), one entry per read; the fields of the reference's
 * vote_extend_kernel outputs, linear_gpu_locate.mojo:778-787. */
typedef struct {
    int32_t *cid;
    uint32_t *pos;
    uint32_t *mapq;
    int32_t *flag;
    uint32_t *sl;
    uint32_t *sr;
    uint32_t *nm;
    uint32_t *iop;
    uint32_t *iat;
    uint32_t *ilen;
} or_recs;

/* Algorithmic-traffic counters (per batch). */
typedef struct {
    int64_t probes;    /* valid keys looked up */
    int64_t postings;  /* postings read (vote + rescue + high-occ tally) */
    int64_t windows;   /* reference windows compared (extend + rescue + tally) */
    int64_t reads_rc;  /* reads that ran the RC pass */
    int64_t overflow;  /* vote-table first-min replacements attempted (table full) */
    int64_t tally;     /* reads placed by the high-occ exact tally */
} or_counters;

static void rec_reset(or_recs *o, int64_t r) {
    o->cid[r] = -1;
    o->pos[r] = 0;
    o->mapq[r] = 0;
    o->flag[r] = 4;
    o->sl[r] = 0;
    o->sr[r] = 0;
    o->nm[r] = 0;
    o->iop[r] = 0;
    o->iat[r] = 0;
    o->ilen[r] = 0;
}

static inline int mism(uint8_t a, uint8_t b) { return a > 3 || b > 3 || a != b; }

#define OR_TOP 256
#define OR_KEEP 16
#define OR_MAX_SLOTS 128
#define OR_HX 48

/* Per-read vote -> KEEP -> gapless/softclip(/indel) extend -> rescue ->
 * high-occ exact tally -> MAPQ. pass_mode 0 = vote + extend + low-occ
 * rescue; pass_mode 1 = high-occ exact tally only (read must be unmapped).
 * Follows vote_extend_kernel, linear_gpu_locate.mojo:749-1607 (pass_mode 2 is
 * dead code there, :863-902, and is not restated).
 * Counter increments are approximate for traffic accounting only. */
static void vote_extend_read(const or_index *ix, const uint8_t *codes, const uint32_t *lens,
                             int64_t max_len, int64_t slots_per, const uint64_t *occ_start,
                             const uint64_t *occ_end, const uint32_t *q_offs,
                             int64_t max_diff, int64_t max_soft, int64_t max_indel,
                             int64_t vote_occ, int pass_mode, int32_t strand_flag, int64_t rid,
                             or_recs *out, int64_t *c_post, int64_t *c_win, int64_t *c_ovf,
                             int64_t *c_tally) {
    /* :800-802 pass-1 safety */
    if (pass_mode == 1 && out->cid[rid] >= 0) return;
    rec_reset(out, rid);
    int64_t qlen = (int64_t)lens[rid];
    if (qlen <= 0) return;

    uint64_t cand_key[OR_TOP];
    int32_t cand_n[OR_TOP];
    int64_t n_cand = 0;

    /* :819-853 valid slots ordered by ascending occ (stable insertion). */
    int32_t ord_slot[OR_MAX_SLOTS];
    int32_t ord_occ[OR_MAX_SLOTS];
    int64_t n_ord = 0;
    for (int64_t sg = 0; sg < slots_per && n_ord < OR_MAX_SLOTS; sg++) {
        int64_t si = rid * slots_per + sg;
        int64_t a = (int64_t)occ_start[si], b = (int64_t)occ_end[si];
        int64_t occ = b - a;
        if (occ > 0 && a >= 0 && b <= ix->n_postings) {
            int64_t ins = n_ord;
            for (int64_t t = 0; t < n_ord; t++)
                if (occ < (int64_t)ord_occ[t]) { ins = t; break; }
            for (int64_t s = n_ord; s > ins; s--) {
                ord_slot[s] = ord_slot[s - 1];
                ord_occ[s] = ord_occ[s - 1];
            }
            ord_slot[ins] = (int32_t)sg;
            ord_occ[ins] = (int32_t)occ;
            n_ord++;
        }
    }
    int64_t vote_cap = vote_occ < 1 ? 256 : vote_occ;

    /* :903-973 serial vote into TOP with first-min replacement and early stop. */
    if (pass_mode == 0) {
        int64_t last_i = -1;
        for (int64_t oi = 0; oi < n_ord; oi++) {
            int64_t occ_v = ord_occ[oi];
            if (occ_v > vote_cap) continue;
            int64_t si = rid * slots_per + ord_slot[oi];
            int64_t a = (int64_t)occ_start[si], b = (int64_t)occ_end[si];
            int64_t q_off = (int64_t)q_offs[si];
            for (int64_t pi = a; pi < b; pi++) {
                int64_t cid = (int64_t)ix->postings[pi * 2];
                int64_t pos = (int64_t)ix->postings[pi * 2 + 1];
                (*c_post)++;
                if (!(cid >= 0 && cid < ix->n_contigs && pos >= q_off)) continue;
                uint64_t vk = ((uint64_t)cid << 32) | (uint64_t)(pos - q_off);
                int64_t found = -1;
                if (last_i >= 0 && cand_key[last_i] == vk) {
                    found = last_i;
                } else {
                    for (int64_t ci = 0; ci < n_cand; ci++)
                        if (cand_key[ci] == vk) { found = ci; break; }
                }
                if (found >= 0) {
                    cand_n[found]++;
                    last_i = found;
                } else if (n_cand < OR_TOP) {
                    cand_key[n_cand] = vk;
                    cand_n[n_cand] = 1;
                    last_i = n_cand;
                    n_cand++;
                } else {
                    (*c_ovf)++;
                    int64_t min_i = 0;
                    int32_t min_n = cand_n[0];
                    for (int64_t z = 1; z < OR_TOP; z++)
                        if (cand_n[z] < min_n) { min_n = cand_n[z]; min_i = z; }
                    if (min_n <= 1) {
                        cand_key[min_i] = vk;
                        cand_n[min_i] = 1;
                        last_i = min_i;
                    }
                }
            }
            if (n_cand > 0 && oi + 1 < n_ord) {
                int64_t b1 = 0, b2 = 0;
                for (int64_t z = 0; z < n_cand; z++) {
                    int64_t zv = cand_n[z];
                    if (zv > b1) { b2 = b1; b1 = zv; }
                    else if (zv > b2) b2 = zv;
                }
                if (b1 >= 4 && b1 >= b2 + 2) break;
            }
        }
    }

    /* :975-1004 stable top-KEEP by votes (strict >). */
    int32_t pick_i[OR_KEEP];
    int32_t pick_n[OR_KEEP];
    int64_t n_pick = 0;
    if (pass_mode == 0) {
        for (int64_t c = 0; c < n_cand; c++) {
            int32_t votes = cand_n[c];
            int inserted = 0;
            for (int64_t p = 0; p < n_pick; p++) {
                if (votes > pick_n[p]) {
                    int64_t s = n_pick;
                    if (s > OR_KEEP - 1) s = OR_KEEP - 1;
                    for (; s > p; s--) {
                        pick_i[s] = pick_i[s - 1];
                        pick_n[s] = pick_n[s - 1];
                    }
                    pick_i[p] = (int32_t)c;
                    pick_n[p] = votes;
                    if (n_pick < OR_KEEP) n_pick++;
                    inserted = 1;
                    break;
                }
            }
            if (!inserted && n_pick < OR_KEEP) {
                pick_i[n_pick] = (int32_t)c;
                pick_n[n_pick] = votes;
                n_pick++;
            }
        }
    }

    /* :1006-1024 caps. */
    const uint8_t *q = codes + rid * max_len;
    const uint8_t *R = ix->ref_codes;
    int64_t budget = max_diff;
    if (budget < 2) budget = 2;
    if (budget > qlen / 2) budget = qlen / 2;
    int64_t clip_cap = max_soft;
    if (clip_cap > qlen / 4) clip_cap = qlen / 4;
    int64_t indel_cap = max_indel;
    if (indel_cap < 0) indel_cap = 0;
    if (indel_cap > 4) indel_cap = 4;
    int64_t shift_cap = indel_cap;
    if (shift_cap < 2) shift_cap = 2;
    if (shift_cap > 6) shift_cap = 6;

    int64_t best_cost = 9999;
    int32_t best_votes = 0;
    int64_t best_cid = -1, best_pos = 0, best_sl = 0, best_sr = 0;
    int64_t best_iop = 0, best_iat = 0, best_ilen = 0;

    /* :1037-1243 extend picks at adj 0, -1, +1, -2, +2, ... */
    for (int64_t pk = 0; pk < n_pick && n_cand > 0; pk++) {
        uint64_t bkey = cand_key[pick_i[pk]];
        int32_t votes = pick_n[pk];
        int64_t bcid = (int64_t)(bkey >> 32);
        int64_t bstart0 = (int64_t)(bkey & 0xFFFFFFFFull);
        if (bcid >= 0 && bcid < ix->n_contigs && bstart0 >= 0) {
            int64_t off0 = (int64_t)ix->contig_off[bcid];
            int64_t clen = (int64_t)ix->contig_off[bcid + 1] - off0;
            for (int64_t adj_i = 0; adj_i <= shift_cap * 2; adj_i++) {
                int64_t adj = 0;
                if (adj_i > 0) {
                    int64_t mag = (adj_i + 1) / 2;
                    adj = (adj_i & 1) ? -mag : mag;
                }
                int64_t bstart = bstart0 + adj;
                if (bstart < 0 || bstart + qlen > clen) continue;
                const uint8_t *r = R + off0 + bstart;
                (*c_win)++;
                int64_t probe_nm = 0;
                for (int64_t pkp = 1; pkp <= 6; pkp++) {
                    int64_t jp = (pkp * (qlen - 1)) / 7;
                    if (jp < clip_cap) jp = clip_cap;
                    if (jp > qlen - 1 - clip_cap) jp = qlen - 1 - clip_cap;
                    if (jp >= 0 && jp < qlen && mism(r[jp], q[jp])) probe_nm++;
                }
                if (probe_nm > budget + indel_cap) continue;
                int64_t nm_all = 0;
                for (int64_t j = 0; j < qlen && nm_all <= budget; j++)
                    if (mism(r[j], q[j])) nm_all++;
                int64_t sl = 0, sr = 0, nm = nm_all;
                if (nm_all > budget) {
                    while (sl < clip_cap) {
                        if (!mism(r[sl], q[sl])) break;
                        sl++;
                    }
                    while (sr < clip_cap) {
                        int64_t jr = qlen - 1 - sr;
                        if (jr <= sl) break;
                        if (!mism(r[jr], q[jr])) break;
                        sr++;
                    }
                    nm = 0;
                    for (int64_t j = sl; j < qlen - sr && nm <= budget; j++)
                        if (mism(r[j], q[j])) nm++;
                }
                int64_t alen = qlen - sl - sr;
                if (nm <= budget && alen >= 24) {
                    if (nm < best_cost || (nm == best_cost && votes > best_votes)) {
                        best_cost = nm;
                        best_votes = votes;
                        best_cid = bcid;
                        best_pos = bstart + sl;
                        best_sl = sl;
                        best_sr = sr;
                        best_iop = best_iat = best_ilen = 0;
                    }
                }
                /* :1142-1235 single indel at adj 0 on a coarse grid (step 2). */
                if (indel_cap > 0 && adj == 0 && (nm > budget || alen < 24)) {
                    for (int64_t dlen = 1; dlen <= indel_cap; dlen++) {
                        if (bstart + qlen + dlen <= clen) {
                            for (int64_t g = 0; g <= qlen; g += 2) {
                                int64_t nd = dlen, t = 0;
                                for (; t < g && nd <= budget; t++)
                                    if (mism(r[t], q[t])) nd++;
                                for (; t < qlen && nd <= budget; t++)
                                    if (mism(r[t + dlen], q[t])) nd++;
                                if (nd <= budget &&
                                    (nd < best_cost || (nd == best_cost && votes > best_votes))) {
                                    best_cost = nd;
                                    best_votes = votes;
                                    best_cid = bcid;
                                    best_pos = bstart;
                                    best_sl = best_sr = 0;
                                    best_iop = 2;
                                    best_iat = g;
                                    best_ilen = dlen;
                                }
                            }
                        }
                        if (qlen > dlen && bstart + (qlen - dlen) <= clen) {
                            for (int64_t gi = 0; gi <= qlen - dlen; gi += 2) {
                                int64_t ni = dlen, ti = 0;
                                for (; ti < gi && ni <= budget; ti++)
                                    if (mism(r[ti], q[ti])) ni++;
                                for (int64_t tj = gi + dlen, rj = gi; tj < qlen && ni <= budget; tj++, rj++)
                                    if (mism(r[rj], q[tj])) ni++;
                                if (ni <= budget &&
                                    (ni < best_cost || (ni == best_cost && votes > best_votes))) {
                                    best_cost = ni;
                                    best_votes = votes;
                                    best_cid = bcid;
                                    best_pos = bstart;
                                    best_sl = best_sr = 0;
                                    best_iop = 1;
                                    best_iat = gi;
                                    best_ilen = dlen;
                                }
                            }
                        }
                    }
                }
                if (best_cost == 0) break;
            }
            if (best_cost == 0) break;
        }
    }

    /* :1245-1532 low-occ seed rescue (pass 0, only if unmapped) and high-occ
     * exact tally (pass 1). */
    if (pass_mode == 1 || best_cid < 0) {
        uint64_t hx_key[OR_HX];
        int32_t hx_n[OR_HX];
        int32_t hx_min_occ[OR_HX];
        int64_t n_hx = 0, n_high_tried = 0;
        int64_t ri = 0;
        while (ri < n_ord) {
            int64_t occ_r = ord_occ[ri];
            if (occ_r <= 0) { ri++; continue; }
            int high = occ_r > vote_cap;
            if (high && pass_mode == 0) { ri++; continue; }
            if (!high && pass_mode == 1) { ri++; continue; }
            if (high) {
                n_high_tried++;
                if (n_high_tried > 96) { ri++; continue; }
            }
            int64_t si = rid * slots_per + ord_slot[ri];
            int64_t a_r = (int64_t)occ_start[si], b_r = (int64_t)occ_end[si];
            int64_t q_off_r = (int64_t)q_offs[si];
            if (high) {
                for (int64_t ph = a_r; ph < b_r; ph++) {
                    int64_t cid_h = (int64_t)ix->postings[ph * 2];
                    int64_t pos_h = (int64_t)ix->postings[ph * 2 + 1];
                    (*c_post)++;
                    if (!(cid_h >= 0 && cid_h < ix->n_contigs && pos_h >= q_off_r)) continue;
                    int64_t bs_h = pos_h - q_off_r;
                    int64_t o0 = (int64_t)ix->contig_off[cid_h];
                    int64_t o1 = (int64_t)ix->contig_off[cid_h + 1];
                    if (!(bs_h >= 0 && bs_h + qlen <= o1 - o0)) continue;
                    const uint8_t *r = R + o0 + bs_h;
                    (*c_win)++;
                    int exact = 1;
                    for (int64_t j = 0; j < qlen; j++)
                        if (mism(r[j], q[j])) { exact = 0; break; }
                    if (!exact) continue;
                    uint64_t vk = ((uint64_t)cid_h << 32) | (uint64_t)bs_h;
                    int64_t f = -1;
                    for (int64_t c = 0; c < n_hx; c++)
                        if (hx_key[c] == vk) { f = c; break; }
                    if (f >= 0) {
                        hx_n[f]++;
                        if ((int32_t)occ_r < hx_min_occ[f]) hx_min_occ[f] = (int32_t)occ_r;
                    } else if (n_hx < OR_HX) {
                        hx_key[n_hx] = vk;
                        hx_n[n_hx] = 1;
                        hx_min_occ[n_hx] = (int32_t)occ_r;
                        n_hx++;
                    }
                }
                if (n_hx > 0) {
                    int64_t b1 = 0, b2 = 0;
                    for (int64_t z = 0; z < n_hx; z++) {
                        int64_t zv = hx_n[z];
                        if (zv > b1) { b2 = b1; b1 = zv; }
                        else if (zv > b2) b2 = zv;
                    }
                    if (b1 >= 3 && b1 > b2)
                        while (ri + 1 < n_ord && (int64_t)ord_occ[ri + 1] > vote_cap) ri++;
                }
                ri++;
                continue;
            }
            /* low-occ rescue: gapless -> softclip -> single indel near seed */
            for (int64_t pr = a_r; pr < b_r; pr++) {
                int64_t cid_r = (int64_t)ix->postings[pr * 2];
                int64_t pos_r = (int64_t)ix->postings[pr * 2 + 1];
                (*c_post)++;
                if (!(cid_r >= 0 && cid_r < ix->n_contigs && pos_r >= q_off_r)) continue;
                int64_t bstart_r = pos_r - q_off_r;
                int64_t off0_r = (int64_t)ix->contig_off[cid_r];
                int64_t clen_r = (int64_t)ix->contig_off[cid_r + 1] - off0_r;
                if (!(bstart_r >= 0 && bstart_r + qlen <= clen_r)) continue;
                const uint8_t *r = R + off0_r + bstart_r;
                (*c_win)++;
                int64_t nm_r = 0;
                for (int64_t j = 0; j < qlen && nm_r <= budget; j++)
                    if (mism(r[j], q[j])) nm_r++;
                int64_t sl_r = 0, sr_r = 0, cost_r = nm_r;
                int64_t iop_r = 0, iat_r = 0, ilen_r = 0, pos_out = bstart_r;
                if (nm_r > budget && clip_cap > 0) {
                    while (sl_r < clip_cap) {
                        if (!mism(r[sl_r], q[sl_r])) break;
                        sl_r++;
                    }
                    while (sr_r < clip_cap) {
                        int64_t jr = qlen - 1 - sr_r;
                        if (jr <= sl_r) break;
                        if (!mism(r[jr], q[jr])) break;
                        sr_r++;
                    }
                    cost_r = 0;
                    for (int64_t j = sl_r; j < qlen - sr_r && cost_r <= budget; j++)
                        if (mism(r[j], q[j])) cost_r++;
                    pos_out = bstart_r + sl_r;
                    if (qlen - sl_r - sr_r < 32) cost_r = 9999;
                }
                if (indel_cap > 0 && cost_r > budget) {
                    for (int64_t dlen = 1; dlen <= indel_cap; dlen++) {
                        if (bstart_r + qlen + dlen <= clen_r) {
                            int64_t g = q_off_r;
                            if (g > qlen) g = qlen;
                            int64_t g0 = g - 8 < 0 ? 0 : g - 8;
                            int64_t g1 = g + 8 > qlen ? qlen : g + 8;
                            for (int64_t gg = g0; gg <= g1; gg++) {
                                int64_t nd = dlen, t = 0;
                                for (; t < gg && nd <= budget; t++)
                                    if (mism(r[t], q[t])) nd++;
                                for (; t < qlen && nd <= budget; t++)
                                    if (mism(r[t + dlen], q[t])) nd++;
                                if (nd < cost_r) {
                                    cost_r = nd;
                                    sl_r = sr_r = 0;
                                    iop_r = 2;
                                    iat_r = gg;
                                    ilen_r = dlen;
                                    pos_out = bstart_r;
                                }
                            }
                        }
                        if (qlen > dlen && bstart_r + (qlen - dlen) <= clen_r) {
                            int64_t g2 = q_off_r;
                            if (g2 > qlen - dlen) g2 = qlen - dlen;
                            int64_t g0i = g2 - 8 < 0 ? 0 : g2 - 8;
                            int64_t g1i = g2 + 8 > qlen - dlen ? qlen - dlen : g2 + 8;
                            for (int64_t gi = g0i; gi <= g1i; gi++) {
                                int64_t ni = dlen, ti = 0;
                                for (; ti < gi && ni <= budget; ti++)
                                    if (mism(r[ti], q[ti])) ni++;
                                for (int64_t tj = gi + dlen, rj = gi; tj < qlen && ni <= budget; tj++, rj++)
                                    if (mism(r[rj], q[tj])) ni++;
                                if (ni < cost_r) {
                                    cost_r = ni;
                                    sl_r = sr_r = 0;
                                    iop_r = 1;
                                    iat_r = gi;
                                    ilen_r = dlen;
                                    pos_out = bstart_r;
                                }
                            }
                        }
                    }
                }
                if (cost_r <= budget && (cost_r < best_cost || best_cid < 0)) {
                    best_cost = cost_r;
                    best_votes = occ_r <= 64 ? 5 : 4;
                    best_cid = cid_r;
                    best_pos = pos_out;
                    best_sl = sl_r;
                    best_sr = sr_r;
                    best_iop = iop_r;
                    best_iat = iat_r;
                    best_ilen = ilen_r;
                    if (cost_r == 0 && iop_r == 0) {
                        pr = b_r; /* loop increment leaves pr > b_r: ends the scan */
                        ri = n_ord;
                    }
                }
            }
            if (best_cost == 0 && best_iop == 0 && best_cid >= 0) break;
            ri++;
        }

        /* :1534-1585 resolve the high-occ exact tally. */
        if (n_hx > 0) {
            int64_t best_hi = 0, second_hi = 0, best_hx_i = -1, best_mocc = 0x7FFFFFFF;
            for (int64_t h = 0; h < n_hx; h++) {
                int64_t vn = hx_n[h], mo = hx_min_occ[h];
                if (vn > best_hi) {
                    second_hi = best_hi;
                    best_hi = vn;
                    best_mocc = mo;
                    best_hx_i = h;
                } else if (vn == best_hi && mo < best_mocc) {
                    best_mocc = mo;
                    best_hx_i = h;
                } else if (vn > second_hi) {
                    second_hi = vn;
                }
            }
            if (best_hx_i >= 0 && (best_hi >= 2 || (best_hi >= 1 && second_hi == 0)) &&
                (best_cid < 0 || best_cost > 0 || (int64_t)best_votes < best_hi)) {
                uint64_t bk = hx_key[best_hx_i];
                (*c_tally)++;
                best_cost = 0;
                if (best_hi > second_hi) {
                    best_votes = (int32_t)(best_hi > 5 ? 5 : best_hi);
                } else {
                    best_votes = 1;
                }
                best_cid = (int64_t)(bk >> 32);
                best_pos = (int64_t)(bk & 0xFFFFFFFFull);
                best_sl = best_sr = 0;
                best_iop = best_iat = best_ilen = 0;
            }
        }
    }

    /* :1587-1607 MAPQ and outputs. */
    if (best_cid >= 0) {
        uint32_t mq = 20;
        if (best_votes >= 3) mq = 40;
        if (best_votes >= 5) mq = 60;
        if (best_cost > 2 && mq > 20) mq = 20;
        if (best_iop != 0 && mq > 20) mq = 20;
        out->cid[rid] = (int32_t)best_cid;
        out->pos[rid] = (uint32_t)best_pos;
        out->mapq[rid] = mq;
        out->flag[rid] = strand_flag;
        out->sl[rid] = (uint32_t)best_sl;
        out->sr[rid] = (uint32_t)best_sr;
        out->nm[rid] = (uint32_t)best_cost;
        out->iop[rid] = (uint32_t)best_iop;
        out->iat[rid] = (uint32_t)best_iat;
        out->ilen[rid] = (uint32_t)best_ilen;
    }
}

/* Kernel-level entry: run vote_extend over reads (n_active < 0: all n_reads;
 * else the rid_map list). Exposed for branch-level tests. */
void or_vote_extend(const or_index *ix, const uint8_t *codes, const uint32_t *lens,
                    int64_t max_len, int64_t slots_per, int64_t n_reads,
                    const uint64_t *occ_start, const uint64_t *occ_end, const uint32_t *q_offs,
                    int64_t max_diff, int64_t max_soft, int64_t max_indel, int64_t vote_occ,
                    int pass_mode, int64_t n_active, const int32_t *rid_map, int32_t strand_flag,
                    or_recs *out) {
    int64_t n = n_active >= 0 ? n_active : n_reads;
    int64_t cp = 0, cw = 0, co = 0, ct = 0;
    for (int64_t t = 0; t < n; t++) {
        int64_t rid = n_active >= 0 ? rid_map[t] : t;
        vote_extend_read(ix, codes, lens, max_len, slots_per, occ_start, occ_end, q_offs,
                         max_diff, max_soft, max_indel, vote_occ, pass_mode, strand_flag, rid,
                         out, &cp, &cw, &co, &ct);
    }
}

/* ------------------------------------------------------ A5 + A8 driver */

static int64_t count_valid_keys(const uint64_t *keys, int64_t from, int64_t n) {
    int64_t c = 0;
    for (int64_t i = from; i < from + n; i++) c += keys[i] != OR_EMPTY_KEY;
    return c;
}

/* One read through the whole per-batch sequence of the reference's host
 * driver: FW encode/locate/pass 0 (vote_fast) -> optional pass 0 at vote_occ ->
 * pass 1 if unmapped -> RC (need_rc = FW unmapped or FW nm > 0) with the same
 * passes -> FW/RC merge. Per-read results do not depend on the batch.
 * Follows linear_gpu_locate.mojo:1936-2032 (FW), 2215-2342 (pass 1),
 * 2344-2635 (RC), 2101-2113 (merge), 548-600 (RC codes), 217-240 (hit). */
static void map_one_read(const or_index *ix, const or_params *P, const uint8_t *codes_all,
                         const uint32_t *lens, int64_t max_len, int64_t slots_per, int64_t rid,
                         uint8_t *rc_buf, uint64_t *keys, uint32_t *qoffs, uint64_t *st,
                         uint64_t *en, or_recs *fw, or_recs *rc, or_recs *mg, or_counters *cnt) {
    int64_t base = rid * slots_per;
    int64_t cp = 0, cw = 0, co = 0, ct = 0;
    /* FW */
    encode_one_read(codes_all, rid, lens[rid], max_len, P->k, P->stride, slots_per, keys, qoffs);
    for (int64_t s = 0; s < slots_per; s++)
        locate_one(keys[base + s], ix->kmers, ix->offsets, ix->n_keys, P->max_occ, &st[base + s],
                   &en[base + s]);
    cnt->probes += count_valid_keys(keys, base, slots_per);
    vote_extend_read(ix, codes_all, lens, max_len, slots_per, st, en, qoffs, P->max_diff,
                     P->max_soft, P->max_indel, P->vote_fast, 0, 0, rid, fw, &cp, &cw, &co, &ct);
    if (fw->cid[rid] < 0) {
        if (P->vote_fast < P->vote_occ)
            vote_extend_read(ix, codes_all, lens, max_len, slots_per, st, en, qoffs, P->max_diff,
                             P->max_soft, P->max_indel, P->vote_occ, 0, 0, rid, fw, &cp, &cw, &co, &ct);
        if (fw->cid[rid] < 0)
            vote_extend_read(ix, codes_all, lens, max_len, slots_per, st, en, qoffs, P->max_diff,
                             P->max_soft, P->max_indel, P->vote_occ, 1, 0, rid, fw, &cp, &cw, &co, &ct);
    }
    /* RC */
    rec_reset(rc, rid);
    if (P->do_rc && (fw->cid[rid] < 0 || fw->nm[rid] > 0)) {
        cnt->reads_rc++;
        int64_t L = (int64_t)lens[rid];
        const uint8_t *src = codes_all + rid * max_len;
        uint8_t *dst = rc_buf + rid * max_len;
        for (int64_t i = 0; i < max_len; i++) dst[i] = 255;
        for (int64_t j = 0; j < L; j++) {
            uint8_t c = src[L - 1 - j];
            dst[j] = c <= 3 ? (uint8_t)(c ^ 3) : 255;
        }
        encode_one_read(rc_buf, rid, lens[rid], max_len, P->k, P->stride, slots_per, keys, qoffs);
        for (int64_t s = 0; s < slots_per; s++)
            locate_one(keys[base + s], ix->kmers, ix->offsets, ix->n_keys, P->max_occ,
                       &st[base + s], &en[base + s]);
        cnt->probes += count_valid_keys(keys, base, slots_per);
        vote_extend_read(ix, rc_buf, lens, max_len, slots_per, st, en, qoffs, P->max_diff,
                         P->max_soft, P->max_indel, P->vote_fast, 0, 16, rid, rc, &cp, &cw, &co, &ct);
        if (rc->cid[rid] < 0) {
            if (P->vote_fast < P->vote_occ)
                vote_extend_read(ix, rc_buf, lens, max_len, slots_per, st, en, qoffs,
                                 P->max_diff, P->max_soft, P->max_indel, P->vote_occ, 0, 16, rid,
                                 rc, &cp, &cw, &co, &ct);
            if (rc->cid[rid] < 0)
                vote_extend_read(ix, rc_buf, lens, max_len, slots_per, st, en, qoffs,
                                 P->max_diff, P->max_soft, P->max_indel, P->vote_occ, 1, 16, rid,
                                 rc, &cp, &cw, &co, &ct);
        }
    }
    /* merge: RC if FW unmapped and RC mapped, or both mapped and RC has lower
     * nm (tie -> higher mapq). */
    int use_rc = 0;
    if (fw->cid[rid] < 0 && rc->cid[rid] >= 0) use_rc = 1;
    else if (fw->cid[rid] >= 0 && rc->cid[rid] >= 0) {
        if (rc->nm[rid] < fw->nm[rid] || (rc->nm[rid] == fw->nm[rid] && rc->mapq[rid] > fw->mapq[rid]))
            use_rc = 1;
    }
    const or_recs *s = use_rc ? rc : fw;
    mg->cid[rid] = s->cid[rid];
    mg->pos[rid] = s->pos[rid];
    mg->mapq[rid] = s->mapq[rid];
    mg->flag[rid] = s->flag[rid];
    mg->sl[rid] = s->sl[rid];
    mg->sr[rid] = s->sr[rid];
    mg->nm[rid] = s->nm[rid];
    mg->iop[rid] = s->iop[rid];
    mg->iat[rid] = s->iat[rid];
    mg->ilen[rid] = s->ilen[rid];
    /* _hit_from_gpu (:229-234): unmapped, or no aligned middle -> unmapped. */
    int64_t m = (int64_t)lens[rid] - (int64_t)mg->sl[rid] - (int64_t)mg->sr[rid];
    if (mg->cid[rid] < 0 || m < 1) rec_reset(mg, rid);
    cnt->postings += cp;
    cnt->windows += cw;
    cnt->overflow += co;
    cnt->tally += ct;
}

/* Slots per read for a batch: S = (max(max_len, k) - k) / stride + 1
 * (linear_gpu_locate.mojo:1853-1857). */
int64_t or_slots_per(int64_t max_len, int64_t k, int64_t stride) {
    if (max_len < k) max_len = k;
    return (max_len - k) / stride + 1;
}

/* Full batch over reads in the reference's padded layout: bases[n x max_len]
 * ASCII (already in-silico converted), padded with 'N'. Outputs FW, RC and
 * merged records. Parallel over reads (OpenMP) when nthreads > 1. */
int or_map_batch(const or_index *ix, const or_params *P, const uint8_t *bases,
                 const uint32_t *lens, int64_t n, int64_t max_len, or_recs *fw, or_recs *rc,
                 or_recs *mg, or_counters *cnt, int nthreads) {
    if (max_len < P->k) return -1; /* caller pads to >= k like :1853 */
    int64_t S = or_slots_per(max_len, P->k, P->stride);
    int64_t nb = n * max_len;
    uint8_t *codes = (uint8_t *)malloc((size_t)(nb > 0 ? nb : 1));
    uint8_t *rcb = (uint8_t *)malloc((size_t)(nb > 0 ? nb : 1));
    uint64_t *keys = (uint64_t *)malloc((size_t)(n * S > 0 ? n * S : 1) * 8);
    uint32_t *qoffs = (uint32_t *)malloc((size_t)(n * S > 0 ? n * S : 1) * 4);
    uint64_t *st = (uint64_t *)malloc((size_t)(n * S > 0 ? n * S : 1) * 8);
    uint64_t *en = (uint64_t *)malloc((size_t)(n * S > 0 ? n * S : 1) * 8);
    if (!codes || !rcb || !keys || !qoffs || !st || !en) {
        free(codes); free(rcb); free(keys); free(qoffs); free(st); free(en);
        return -2;
    }
    or_counters total = {0, 0, 0, 0, 0, 0};
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        or_counters local = {0, 0, 0, 0, 0, 0};
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t r = 0; r < n; r++) {
            or_pack_bases(bases + r * max_len, codes + r * max_len, max_len);
            map_one_read(ix, P, codes, lens, max_len, S, r, rcb, keys, qoffs, st, en, fw, rc, mg,
                         &local);
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            total.probes += local.probes;
            total.postings += local.postings;
            total.windows += local.windows;
            total.reads_rc += local.reads_rc;
            total.overflow += local.overflow;
            total.tally += local.tally;
        }
    }
    if (cnt) *cnt = total;
    free(codes); free(rcb); free(keys); free(qoffs); free(st); free(en);
    return 0;
}

/* ------------------------------------------------------------- A10 */

/* Follows _unclipped5, linear_gpu_fm.mojo:371-386 (reverse reads use sl). */
int64_t or_unclipped5(int64_t flag, int64_t pos0, int64_t sl, int64_t sr, int64_t qlen) {
    if ((flag & 4) != 0 || qlen <= 0) return 0;
    if (sl < 0) sl = 0;
    if (sr < 0) sr = 0;
    if (sl + sr >= qlen) { sl = 0; sr = 0; }
    int64_t mid = qlen - sl - sr;
    if ((flag & 16) != 0) return pos0 + mid + sl - 1;
    return pos0 - sl;
}

/* Follows _coord_key, linear_gpu_fm.mojo:389-395. */
uint64_t or_coord_key(int64_t tid, int64_t pos0, int64_t orig) {
    if (tid < 0) return (0xFFFFFFFFull << 32) | ((uint64_t)orig & 0xFFFFFFFFull);
    if (pos0 < 0) pos0 = 0;
    return ((uint64_t)tid << 32) | (uint64_t)pos0;
}

/* Follows _pack_end, linear_gpu_fm.mojo:398-405. */
uint64_t or_pack_end(int64_t tid, int64_t u5, int64_t strand) {
    uint64_t t = tid >= 0 ? (uint64_t)(tid + 1) : 0;
    int64_t u = u5 + 1073741824;
    if (u < 0) u = 0;
    return (t << 48) | ((uint64_t)(strand & 1) << 47) | ((uint64_t)u & 140737488355327ull);
}

/* Single-end sort keys per merged record. tid_of_cid[n_tid] maps index contig
 * ids to canonical (f/r-stripped) target ids; a contig outside the table has no
 * target (tid -1, unmapped keys), as rid1 < len(rid_to_tid) guards it at
 * linear_gpu_fm.mojo:764-765. pair = pair_base + record index.
 * Follows the SE builder, linear_gpu_fm.mojo:800-823. */
void or_sort_keys_se(int64_t n, const int32_t *cid, const uint32_t *pos, const int32_t *flag,
                     const uint32_t *sl, const uint32_t *sr, const uint32_t *nm,
                     const uint32_t *mapq, const uint32_t *lens, const int32_t *tid_of_cid,
                     int64_t n_tid, int64_t pair_base, uint64_t *coord, uint64_t *dhi, uint64_t *dlo,
                     uint32_t *score, uint32_t *pair) {
    for (int64_t i = 0; i < n; i++) {
        int64_t tid = cid[i] >= 0 && cid[i] < n_tid ? tid_of_cid[cid[i]] : -1;
        int64_t fl = flag[i];
        coord[i] = or_coord_key(tid, (int64_t)pos[i], i);
        if (tid >= 0) {
            dhi[i] = or_pack_end(tid, or_unclipped5(fl, pos[i], sl[i], sr[i], lens[i]), (fl >> 4) & 1);
            dlo[i] = 0;
        } else {
            dhi[i] = ~0ull;
            dlo[i] = (uint64_t)i;
        }
        int64_t nmc = nm[i];
        if (nmc > 255) nmc = 255;
        score[i] = (uint32_t)(((int64_t)mapq[i] << 8) | (255 - nmc));
        pair[i] = (uint32_t)(pair_base + i);
    }
}

/* ------------------------------------------------------ F4c mate rescue */

/* Mate rescue of a paired batch, in place on the merged records (R1 block at
 * [0, n_pairs), R2 block at [n_pairs, 2 n_pairs)); codes are the converted
 * reads' base codes [2 n_pairs x max_len] (0-3, anything else = N), lens their
 * lengths. For a pair with exactly one mate mapped, the other mate is scanned
 * gapless at every start within +-window (clamped to 32..1024) of the mapped
 * mate's position on its contig copy, the strand opposite the mate's first,
 * then the mate's own; a window's mismatches are counted until they pass
 * budget + 4 (budget = clamp(max_diff, 2, len / 2)); the first window of least
 * count within budget wins. With none within budget, a single-base deletion is
 * tried at the first window of least count: split points 0, 4, 8, ... <= len,
 * bases before the split against the window, after it against the window one
 * base on, plus 1; the first least within budget wins. A rescued mate gets the
 * mapped mate's contig, the window start, flag 16 for the reverse strand, no
 * clips, nm, and MAPQ 60 (nm 0), 40 (nm <= 2) or 20. Reads shorter than 24,
 * and windows reaching past the reference's last base, are skipped; a
 * reference N (code > 3) counts as a mismatch. Follows fm_mate_rescue_kernel,
 * fq2bam-meth/src/linear_gpu_fm.mojo:1839-2011. */
static int64_t rescue_mm(const uint8_t *ref, int64_t bstart, const uint8_t *q, int64_t qlen, int rev,
                         int64_t from, int64_t to, int64_t stop) {
    int64_t nm = 0;
    for (int64_t j = from; j < to && nm <= stop; j++) {
        int64_t rb = ref[bstart + j];
        int64_t qc = q[j];
        if (rev) {
            int64_t qr = q[qlen - 1 - j];
            qc = qr <= 3 ? 3 - qr : 4;
        }
        if (rb > 3 || qc > 3 || rb != qc) nm++;
    }
    return nm;
}

void or_mate_rescue(const or_index *ix, const uint8_t *codes, const uint32_t *lens, int64_t max_len,
                    int64_t n_pairs, int64_t window, int64_t max_diff, or_recs *r) {
    const int64_t l_pac = (int64_t)ix->contig_off[ix->n_contigs];
    for (int64_t pi = 0; pi < n_pairs; pi++) {
        const int64_t i1 = pi, i2 = pi + n_pairs;
        const int m1 = r->cid[i1] >= 0, m2 = r->cid[i2] >= 0;
        if (m1 == m2) continue;
        const int64_t src = m1 ? i1 : i2, dst = m1 ? i2 : i1;
        const int64_t crid = r->cid[src];
        if (crid < 0 || crid >= ix->n_contigs) continue;
        const int64_t qlen = lens[dst];
        if (qlen < 24) continue;
        const int64_t coff = (int64_t)ix->contig_off[crid];
        const int64_t clen = (int64_t)ix->contig_off[crid + 1] - coff;
        const int64_t mate_pos = r->pos[src];
        int64_t win = window < 32 ? 32 : (window > 1024 ? 1024 : window);
        int64_t lo = mate_pos - win;
        if (lo < 0) lo = 0;
        int64_t hi = mate_pos + win;
        if (hi > clen - 24) hi = clen - 24;
        if (hi <= lo) continue;
        int64_t budget = max_diff < 2 ? 2 : max_diff;
        if (budget > qlen / 2) budget = qlen / 2;
        const uint8_t *q = codes + dst * max_len;
        int64_t best_nm = 9999, best_pos = 0, best_flag = 0, scan_nm = 9999, scan_local = lo, scan_rev = 0;
        const int mate_rev = (r->flag[src] & 16) != 0;
        for (int pass = 0; pass < 2; pass++) {
            const int is_rev = pass == 0 ? !mate_rev : mate_rev;
            for (int64_t local = lo; local <= hi; local++) {
                const int64_t bstart = coff + local;
                if (bstart < 0 || bstart + qlen > l_pac) continue;
                const int64_t nm = rescue_mm(ix->ref_codes, bstart, q, qlen, is_rev, 0, qlen, budget + 4);
                if (nm < scan_nm) {
                    scan_nm = nm;
                    scan_local = local;
                    scan_rev = is_rev;
                }
                if (nm <= budget && nm < best_nm) {
                    best_nm = nm;
                    best_pos = local;
                    best_flag = is_rev * 16;
                }
            }
        }
        if (best_nm > budget && scan_nm < 9999) {
            const int64_t bstart = coff + scan_local;
            if (bstart >= 0 && bstart + qlen + 1 <= l_pac) {
                for (int64_t g = 0; g <= qlen; g += 4) {
                    int64_t nm_d = 1 + rescue_mm(ix->ref_codes, bstart, q, qlen, (int)scan_rev, 0, g, budget - 1);
                    if (nm_d <= budget)
                        nm_d += rescue_mm(ix->ref_codes, bstart + 1, q, qlen, (int)scan_rev, g, qlen, budget - nm_d);
                    if (nm_d <= budget && nm_d < best_nm) {
                        best_nm = nm_d;
                        best_pos = scan_local;
                        best_flag = scan_rev * 16;
                    }
                }
            }
        }
        if (best_nm <= budget) {
            r->cid[dst] = (int32_t)crid;
            r->pos[dst] = (uint32_t)best_pos;
            r->mapq[dst] = best_nm == 0 ? 60u : (best_nm <= 2 ? 40u : 20u);
            r->flag[dst] = (int32_t)best_flag;
            r->sl[dst] = 0;
            r->sr[dst] = 0;
            r->nm[dst] = (uint32_t)best_nm;
            r->iop[dst] = 0;
            r->iat[dst] = 0;
            r->ilen[dst] = 0;
        }
    }
}

/* ------------------------------------------------------------- A11 / A12 */

typedef struct { uint64_t hi, lo; uint32_t idx; } or_k2;

static void merge_sort_k2(or_k2 *a, or_k2 *tmp, int64_t n) {
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n;
            int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) {
                int take_j = a[j].hi < a[i].hi || (a[j].hi == a[i].hi && a[j].lo < a[i].lo);
                tmp[o++] = take_j ? a[j++] : a[i++];
            }
            while (i < mid) tmp[o++] = a[i++];
            while (j < hi) tmp[o++] = a[j++];
        }
        memcpy(a, tmp, (size_t)n * sizeof(or_k2));
    }
}

/* perm = stable coord order; dup_perm = stable (dhi, dlo) order; is_dup in
 * original order: over each run of equal (dhi, dlo) (dhi = ~0 never a dup),
 * best = max score, ties -> min pair id; every record whose pair differs from
 * the best pair is a duplicate.
 * Follows gpu_sort_markdup, linear_gpu_sort.mojo:24-551 (one tile), and
 * _markdup_runs / order_run_perms, bam_emit.py:577-601, 756-809. */
int or_sort_markdup(int64_t n, const uint64_t *coord, const uint64_t *dhi, const uint64_t *dlo,
                    const uint32_t *score, const uint32_t *pair, uint32_t *is_dup, uint32_t *perm,
                    int do_markdup, uint32_t *dup_perm) {
    if (n <= 0) return 0;
    or_k2 *a = (or_k2 *)malloc((size_t)n * sizeof(or_k2));
    or_k2 *t = (or_k2 *)malloc((size_t)n * sizeof(or_k2));
    if (!a || !t) { free(a); free(t); return -1; }
    for (int64_t i = 0; i < n; i++) is_dup[i] = 0;
    if (do_markdup) {
        for (int64_t i = 0; i < n; i++) { a[i].hi = dhi[i]; a[i].lo = dlo[i]; a[i].idx = (uint32_t)i; }
        merge_sort_k2(a, t, n);
        int64_t gi = 0;
        while (gi < n) {
            uint32_t o0 = a[gi].idx;
            if (dhi[o0] == ~0ull) { is_dup[o0] = 0; gi++; continue; }
            int64_t gj = gi + 1;
            while (gj < n && dhi[a[gj].idx] == dhi[o0] && dlo[a[gj].idx] == dlo[o0]) gj++;
            uint32_t best_pair = pair[o0], best_score = score[o0];
            for (int64_t g = gi; g < gj; g++) {
                uint32_t ok = a[g].idx;
                if (score[ok] > best_score || (score[ok] == best_score && pair[ok] < best_pair)) {
                    best_score = score[ok];
                    best_pair = pair[ok];
                }
            }
            for (int64_t g = gi; g < gj; g++) {
                uint32_t ok = a[g].idx;
                is_dup[ok] = pair[ok] != best_pair ? 1u : 0u;
            }
            gi = gj;
        }
        if (dup_perm)
            for (int64_t i = 0; i < n; i++) dup_perm[i] = a[i].idx;
    } else if (dup_perm) {
        for (int64_t i = 0; i < n; i++) dup_perm[i] = (uint32_t)i;
    }
    for (int64_t i = 0; i < n; i++) { a[i].hi = coord[i]; a[i].lo = 0; a[i].idx = (uint32_t)i; }
    merge_sort_k2(a, t, n);
    for (int64_t i = 0; i < n; i++) perm[i] = a[i].idx;
    free(a);
    free(t);
    return 0;
}

/* Checker for reported records (no reference counterpart): mismatches between
 * each converted read, reverse-complemented when its flag has 16, and the
 * dual reference at (cid, pos) over the aligned bases [sl, L - sr); a base
 * mismatches when either side is not ACGT or the codes differ (the rule of
 * linear_gpu_locate.mojo's gapless extend). -1 for unmapped records, -2 for a
 * window outside its contig. */
void or_recount_nm(const uint8_t *seq, const uint64_t *contig_off, int64_t n_contigs, const uint8_t *reads,
                   int64_t stride, const uint32_t *lens, int64_t n, const int32_t *cid, const uint32_t *pos,
                   const int32_t *flag, const uint32_t *sl, const uint32_t *sr, int32_t *nm_out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        if (cid[i] < 0 || cid[i] >= n_contigs) {
            nm_out[i] = -1;
            continue;
        }
        const int64_t L = lens[i], a = sl[i], b = L - (int64_t)sr[i];
        const int64_t c0 = (int64_t)contig_off[cid[i]], clen = (int64_t)contig_off[cid[i] + 1] - c0;
        if (b < a || (int64_t)pos[i] + (b - a) > clen) {
            nm_out[i] = -2;
            continue;
        }
        const uint8_t *r = reads + i * stride;
        const int rev = (flag[i] & 16) != 0;
        int32_t nm = 0;
        for (int64_t j = a; j < b; j++) {
            uint8_t q = rev ? r[L - 1 - j] : r[j];
            uint8_t cq = or_code_of(q);
            if (rev && cq <= 3) cq = (uint8_t)(3 - cq);
            const uint8_t cr = or_code_of(seq[c0 + (int64_t)pos[i] + (j - a)]);
            nm += (cq > 3 || cr > 3 || cq != cr) ? 1 : 0;
        }
        nm_out[i] = nm;
    }
}

/* ------------------------------------------------ BAM record pack (next row) */

/* Follows _reg2bin, fq2bam-meth/src/linear_gpu_fm.mojo:479-495 (the SAM
 * specification's binning; bam_emit.py:34-48 is the same function). */
int64_t or_reg2bin(int64_t beg, int64_t end) {
    if (end <= beg) end = beg + 1;
    end -= 1;
    if ((beg >> 14) == (end >> 14)) return 4681 + (beg >> 14);
    if ((beg >> 17) == (end >> 17)) return 585 + (beg >> 17);
    if ((beg >> 20) == (end >> 20)) return 73 + (beg >> 20);
    if ((beg >> 23) == (end >> 23)) return 9 + (beg >> 23);
    if ((beg >> 26) == (end >> 26)) return 1 + (beg >> 26);
    return 0;
}

/* Follows _nt16 / _nt16_comp, linear_gpu_fm.mojo:455-477: A C G T (either
 * case) -> 1 2 4 8, anything else 15; the complement swaps 1<->8 and 2<->4. */
uint8_t or_nt16(uint8_t b) {
    if (b == 'A' || b == 'a') return 1;
    if (b == 'C' || b == 'c') return 2;
    if (b == 'G' || b == 'g') return 4;
    if (b == 'T' || b == 't') return 8;
    return 15;
}

uint8_t or_nt16_comp(uint8_t c) {
    if (c == 1) return 8;
    if (c == 8) return 1;
    if (c == 2) return 4;
    if (c == 4) return 2;
    return 15;
}

static void put32(uint8_t *p, int64_t off, int64_t v) {
    uint64_t u = (uint64_t)v;
    p[off] = (uint8_t)(u & 255);
    p[off + 1] = (uint8_t)((u >> 8) & 255);
    p[off + 2] = (uint8_t)((u >> 16) & 255);
    p[off + 3] = (uint8_t)((u >> 24) & 255);
}

/* One BAM alignment record at p + off0 (block_size first); returns the offset
 * after it, or -1 when it would pass cap ("BAM batch buffer overflow").
 * Follows _pack_bam_aln, linear_gpu_fm.mojo:497-665 step by step: clips
 * clamped and dropped when they cover the read, swapped for flag 16; CIGAR
 * slS mM srS only for a mapped record; bin 4680 and tid/pos -1 otherwise;
 * MAPQ and NM clamped to 0..255; SEQ 4-bit, reverse-complemented for flag 16;
 * QUAL - 33 (reversed for flag 16; 255 where < 33) when qual_len == qlen,
 * else 255 throughout; tags RG:Z:<rg> and NM:C. */
int64_t or_pack_bam_aln(uint8_t *p, int64_t off0, int64_t cap, const uint8_t *name,
                        int64_t name_len, int64_t flag, int64_t tid, int64_t pos0, int64_t mapq,
                        int64_t sl, int64_t sr, const uint8_t *seq, int64_t qlen,
                        const uint8_t *qual, int64_t qual_len, int64_t ntid, int64_t npos,
                        int64_t tlen, int64_t nm, const uint8_t *rg, int64_t rg_n) {
    int rev = (flag & 16) != 0;
    int64_t slv = sl < 0 ? 0 : sl, srv = sr < 0 ? 0 : sr;
    if (slv + srv >= qlen) { slv = 0; srv = 0; }
    int64_t mid = qlen - slv - srv;
    if (rev) { int64_t t = slv; slv = srv; srv = t; }
    int mapped = tid >= 0 && (flag & 4) == 0;
    int64_t n_cigar = 0;
    if (mapped) n_cigar = (slv > 0) + (mid > 0) + (srv > 0);
    int64_t l_qname = name_len + 1;
    int64_t block = 32 + l_qname + 4 * n_cigar + (qlen + 1) / 2 + qlen + (3 + rg_n + 1 + 4);
    if (off0 + 4 + block > cap) return -1;
    int64_t bin = 4680, tid_w = -1, pos_w = -1;
    if (mapped) {
        tid_w = tid;
        pos_w = pos0;
        int64_t end = pos0 + mid;
        if (end <= pos0) end = pos0 + 1;
        bin = or_reg2bin(pos0, end);
    }
    if (ntid < 0) { ntid = -1; npos = -1; }
    int64_t mq = mapq < 0 ? 0 : (mapq > 255 ? 255 : mapq);
    put32(p, off0, block);
    put32(p, off0 + 4, tid_w);
    put32(p, off0 + 8, pos_w);
    put32(p, off0 + 12, (int64_t)(uint32_t)((bin << 16) | (mq << 8) | l_qname));
    put32(p, off0 + 16, (int64_t)(uint32_t)((flag << 16) | n_cigar));
    put32(p, off0 + 20, qlen);
    put32(p, off0 + 24, ntid);
    put32(p, off0 + 28, npos);
    put32(p, off0 + 32, tlen);
    int64_t off = off0 + 36;
    for (int64_t i = 0; i < name_len; i++) p[off + i] = name[i];
    p[off + name_len] = 0;
    off += l_qname;
    if (n_cigar > 0) {
        if (slv > 0) { put32(p, off, (slv << 4) | 4); off += 4; }
        if (mid > 0) { put32(p, off, mid << 4); off += 4; }
        if (srv > 0) { put32(p, off, (srv << 4) | 4); off += 4; }
    }
    int64_t si = 0;
    while (si + 1 < qlen) {
        int64_t ia = rev ? qlen - 1 - si : si, ib = rev ? qlen - 2 - si : si + 1;
        uint8_t a = or_nt16(seq[ia]), b = or_nt16(seq[ib]);
        if (rev) { a = or_nt16_comp(a); b = or_nt16_comp(b); }
        p[off++] = (uint8_t)((a << 4) | b);
        si += 2;
    }
    if (si < qlen) {
        uint8_t c = or_nt16(seq[rev ? 0 : si]);
        if (rev) c = or_nt16_comp(c);
        p[off++] = (uint8_t)(c << 4);
    }
    if (qual_len == qlen && qual) {
        for (int64_t i = 0; i < qlen; i++) {
            uint8_t q = qual[rev ? qlen - 1 - i : i];
            p[off + i] = q >= 33 ? (uint8_t)(q - 33) : 255;
        }
    } else {
        for (int64_t i = 0; i < qlen; i++) p[off + i] = 255;
    }
    off += qlen;
    p[off] = 'R'; p[off + 1] = 'G'; p[off + 2] = 'Z';
    off += 3;
    for (int64_t i = 0; i < rg_n; i++) p[off + i] = rg[i];
    p[off + rg_n] = 0;
    off += rg_n + 1;
    p[off] = 'N'; p[off + 1] = 'M'; p[off + 2] = 'C';
    p[off + 3] = (uint8_t)(nm < 0 ? 0 : (nm > 255 ? 255 : nm));
    return off + 4;
}

/* Sets flag 0x400 in the record at p + rec_off; follows _or_dup_flag,
 * linear_gpu_fm.mojo:408-420. */
static void or_dup_flag(uint8_t *p, int64_t rec_off) {
    uint32_t fnc = (uint32_t)p[rec_off + 16] | ((uint32_t)p[rec_off + 17] << 8) |
                   ((uint32_t)p[rec_off + 18] << 16) | ((uint32_t)p[rec_off + 19] << 24);
    uint32_t v = (((fnc >> 16) | 1024u) << 16) | (fnc & 65535u);
    put32(p, rec_off + 16, v);
}

/* Packs n records from per-record fields (int32 arrays; ntid/npos/tlen may be
 * NULL for -1/-1/0) in input order, then gathers them in perm order (NULL =
 * identity) setting 0x400 where apply_dup and is_dup[orig] != 0. Read i: name
 * names[name_off[i] .. name_off[i+1]), bases seq + i*stride (lens[i]), quality
 * qual + i*stride of qual_lens[i] bytes (qual NULL = absent; qual_lens NULL =
 * lens). Returns the stream length or -1 when it would pass cap. Follows the
 * per-batch pack of _emit_bam_from_batch (linear_gpu_fm.mojo:755-800) and
 * _gather_perm_to_writer (:177-225). */
int64_t or_bam_pack_fields(int64_t n, const int32_t *flag, const int32_t *tid, const int32_t *pos,
                           const int32_t *mapq, const int32_t *sl, const int32_t *sr,
                           const int32_t *nm, const int32_t *ntid, const int32_t *npos,
                           const int32_t *tlen, const uint8_t *names, const int64_t *name_off,
                           const uint8_t *seq, int64_t stride, const uint32_t *lens,
                           const uint8_t *qual, const uint32_t *qual_lens, const uint8_t *rg,
                           int64_t rg_n, const uint32_t *perm, const uint32_t *is_dup,
                           int apply_dup, uint8_t *out, int64_t cap) {
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t qlen = lens[i], l_qname = name_off[i + 1] - name_off[i] + 1;
        total += 36 + l_qname + 12 + (qlen + 1) / 2 + qlen + 3 + rg_n + 1 + 4;
    }
    uint8_t *batch = (uint8_t *)malloc((size_t)(total > 0 ? total : 1));
    int64_t *rec = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
    if (!batch || !rec) { free(batch); free(rec); return -1; }
    int64_t off = 0;
    for (int64_t i = 0; i < n; i++) {
        rec[i] = off;
        off = or_pack_bam_aln(batch, off, total, names + name_off[i], name_off[i + 1] - name_off[i],
                              flag[i], tid[i], pos[i], mapq[i], sl[i], sr[i], seq + i * stride,
                              lens[i], qual ? qual + i * stride : NULL,
                              qual ? (qual_lens ? qual_lens[i] : lens[i]) : 0,
                              ntid ? ntid[i] : -1, npos ? npos[i] : -1, tlen ? tlen[i] : 0, nm[i],
                              rg, rg_n);
    }
    rec[n] = off;
    int64_t g = 0;
    for (int64_t j = 0; j < n; j++) {
        int64_t o = perm ? (int64_t)perm[j] : j, len = rec[o + 1] - rec[o];
        if (g + len > cap) { free(batch); free(rec); return -1; }
        memcpy(out + g, batch + rec[o], (size_t)len);
        if (apply_dup && is_dup[o] != 0) or_dup_flag(out, g);
        g += len;
    }
    free(batch);
    free(rec);
    return g;
}

/* Single-end records straight from merged map records: flag | 0x41 (paired +
 * read 1, as the reference's SE branch sets them), tid from tid_of_cid (cid
 * outside [0, n_tid) -> tid -1 and flag | 4), no mate. Follows
 * _emit_bam_from_batch's SE branch, linear_gpu_fm.mojo:755-800. */
int64_t or_bam_pack_se(int64_t n, const int32_t *cid, const uint32_t *pos, const uint32_t *mapq,
                       const int32_t *rflag, const uint32_t *sl, const uint32_t *sr,
                       const uint32_t *nm, const int32_t *tid_of_cid, int64_t n_tid,
                       const uint8_t *names, const int64_t *name_off, const uint8_t *seq,
                       int64_t stride, const uint32_t *lens, const uint8_t *qual,
                       const uint32_t *qual_lens, const uint8_t *rg, int64_t rg_n,
                       const uint32_t *perm, const uint32_t *is_dup, int apply_dup, uint8_t *out,
                       int64_t cap) {
    int32_t *f = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * 7 * sizeof(int32_t));
    if (!f) return -1;
    int32_t *fl = f, *tid = f + n, *ps = f + 2 * n, *mq = f + 3 * n, *l = f + 4 * n, *r = f + 5 * n,
            *m = f + 6 * n;
    for (int64_t i = 0; i < n; i++) {
        int64_t fv = (int64_t)rflag[i] | 1 | 64, t = -1;
        if (cid[i] >= 0 && cid[i] < n_tid) t = tid_of_cid[cid[i]];
        else fv |= 4;
        fl[i] = (int32_t)fv; tid[i] = (int32_t)t; ps[i] = (int32_t)pos[i];
        mq[i] = (int32_t)mapq[i]; l[i] = (int32_t)sl[i]; r[i] = (int32_t)sr[i]; m[i] = (int32_t)nm[i];
    }
    int64_t g = or_bam_pack_fields(n, fl, tid, ps, mq, l, r, m, NULL, NULL, NULL, names, name_off,
                                   seq, stride, lens, qual, qual_lens, rg, rg_n, perm, is_dup,
                                   apply_dup, out, cap);
    free(f);
    return g;
}

/* Paired-end flags, mate fields and sort keys of np pairs whose merged map
 * records lie R1 first (records [0, np)) then R2 ([np, 2 np)); lens are the
 * original read lengths, same order. Output record 2p is pair p's R1 and 2p + 1
 * its R2, the order the reference emits them; row[i] is output record i's
 * record / read row. Any output pointer may be NULL. Follows the PE branch of
 * _emit_bam_from_batch, linear_gpu_fm.mojo:826-870 (flags, mate fields, tlen
 * signed by position order) and :936-977 (coord per record; dhi / dlo, score
 * and pair id shared by the pair; pair id = (pair_base + 2p) / 2). */
void or_pe_fields(int64_t np, const int32_t *cid, const uint32_t *pos, const uint32_t *mapq,
                  const int32_t *rflag, const uint32_t *sl, const uint32_t *sr, const uint32_t *nm,
                  const uint32_t *lens, const int32_t *tid_of_cid, int64_t n_tid, int64_t pair_base,
                  int32_t *o_flag, int32_t *o_tid, int32_t *o_pos, int32_t *o_mapq, int32_t *o_sl,
                  int32_t *o_sr, int32_t *o_nm, int32_t *o_ntid, int32_t *o_npos, int32_t *o_tlen,
                  int32_t *o_row, uint64_t *coord, uint64_t *dhi, uint64_t *dlo, uint32_t *score,
                  uint32_t *pair) {
    for (int64_t p = 0; p < np; p++) {
        int64_t r[2] = {p, np + p}, fl[2], tid[2], ps[2], ntid[2] = {-1, -1}, npos[2] = {-1, -1},
                tlen[2] = {0, 0};
        for (int m = 0; m < 2; m++) {
            int64_t c = cid[r[m]];
            fl[m] = (int64_t)rflag[r[m]] | 1 | (m ? 128 : 64);
            tid[m] = -1;
            if (c >= 0 && c < n_tid) tid[m] = tid_of_cid[c];
            else fl[m] |= 4;
            ps[m] = (int64_t)pos[r[m]];
        }
        if (tid[0] < 0) fl[1] |= 8;
        if (tid[1] < 0) fl[0] |= 8;
        if (tid[0] >= 0 && tid[1] >= 0 && tid[0] == tid[1]) {
            fl[0] |= 2;
            fl[1] |= 2;
            ntid[0] = tid[0];
            ntid[1] = tid[1];
            npos[0] = ps[1];
            npos[1] = ps[0];
            int64_t span = ps[1] - ps[0];
            if (span < 0) span = -span;
            span += (int64_t)lens[r[1]];
            tlen[0] = ps[0] <= ps[1] ? span : -span;
            tlen[1] = -tlen[0];
        } else {
            if (tid[0] >= 0) { ntid[1] = tid[0]; npos[1] = ps[0]; }
            if (tid[1] >= 0) { ntid[0] = tid[1]; npos[0] = ps[1]; }
        }
        uint64_t h = ~0ull, l = (uint64_t)(2 * p);
        if (tid[0] >= 0 || tid[1] >= 0) {
            uint64_t e[2];
            for (int m = 0; m < 2; m++)
                e[m] = or_pack_end(tid[m], or_unclipped5(fl[m], ps[m], sl[r[m]], sr[r[m]], lens[r[m]]),
                                   (fl[m] >> 4) & 1);
            if (tid[0] < 0) { h = e[1]; l = 0; }
            else if (tid[1] < 0) { h = e[0]; l = 0; }
            else if (e[0] <= e[1]) { h = e[0]; l = e[1]; }
            else { h = e[1]; l = e[0]; }
        }
        int64_t nms = (int64_t)nm[r[0]] + (int64_t)nm[r[1]];
        if (nms > 255) nms = 255;
        uint32_t sc = (uint32_t)((((int64_t)mapq[r[0]] + (int64_t)mapq[r[1]]) << 8) | (255 - nms));
        for (int m = 0; m < 2; m++) {
            int64_t i = 2 * p + m;
            if (o_flag) o_flag[i] = (int32_t)fl[m];
            if (o_tid) o_tid[i] = (int32_t)tid[m];
            if (o_pos) o_pos[i] = (int32_t)ps[m];
            if (o_mapq) o_mapq[i] = (int32_t)mapq[r[m]];
            if (o_sl) o_sl[i] = (int32_t)sl[r[m]];
            if (o_sr) o_sr[i] = (int32_t)sr[r[m]];
            if (o_nm) o_nm[i] = (int32_t)nm[r[m]];
            if (o_ntid) o_ntid[i] = (int32_t)ntid[m];
            if (o_npos) o_npos[i] = (int32_t)npos[m];
            if (o_tlen) o_tlen[i] = (int32_t)tlen[m];
            if (o_row) o_row[i] = (int32_t)r[m];
            if (coord) coord[i] = or_coord_key(tid[m], ps[m], i);
            if (dhi) dhi[i] = h;
            if (dlo) dlo[i] = l;
            if (score) score[i] = sc;
            if (pair) pair[i] = (uint32_t)((pair_base + 2 * p) / 2);
        }
    }
}

/* Paired-end BAM records: or_pe_fields, the reads taken in output order
 * (record i reads row row[i]), then or_bam_pack_fields. perm / is_dup index
 * the 2 np output records. */
int64_t or_bam_pack_pe(int64_t np, const int32_t *cid, const uint32_t *pos, const uint32_t *mapq,
                       const int32_t *rflag, const uint32_t *sl, const uint32_t *sr,
                       const uint32_t *nm, const int32_t *tid_of_cid, int64_t n_tid,
                       const uint8_t *names, const int64_t *name_off, const uint8_t *seq,
                       int64_t stride, const uint32_t *lens, const uint8_t *qual,
                       const uint32_t *qual_lens, const uint8_t *rg, int64_t rg_n,
                       const uint32_t *perm, const uint32_t *is_dup, int apply_dup, uint8_t *out,
                       int64_t cap) {
    const int64_t n = 2 * np, nn = n > 0 ? n : 1;
    int32_t *f = (int32_t *)malloc((size_t)nn * 11 * sizeof(int32_t));
    int64_t *noff = (int64_t *)malloc((size_t)(nn + 1) * sizeof(int64_t));
    uint8_t *nb = (uint8_t *)malloc((size_t)(name_off[n] - name_off[0] + 1));
    uint8_t *sq = (uint8_t *)malloc((size_t)(nn * (stride > 0 ? stride : 1)));
    uint8_t *qu = qual ? (uint8_t *)malloc((size_t)(nn * (stride > 0 ? stride : 1))) : NULL;
    uint32_t *ln = (uint32_t *)malloc((size_t)nn * 2 * sizeof(uint32_t));
    int64_t g = -1;
    if (f && noff && nb && sq && ln && (qu || !qual)) {
        int32_t *c[11];
        for (int k = 0; k < 11; k++) c[k] = f + k * n;
        or_pe_fields(np, cid, pos, mapq, rflag, sl, sr, nm, lens, tid_of_cid, n_tid, 0, c[0], c[1], c[2],
                     c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], NULL, NULL, NULL, NULL, NULL);
        uint32_t *ql = ln + n;
        noff[0] = 0;
        for (int64_t i = 0; i < n; i++) {
            int64_t r = c[10][i], a = name_off[r], b = name_off[r + 1];
            memcpy(nb + noff[i], names + a, (size_t)(b - a));
            noff[i + 1] = noff[i] + (b - a);
            memcpy(sq + i * stride, seq + r * stride, (size_t)stride);
            if (qu) memcpy(qu + i * stride, qual + r * stride, (size_t)stride);
            ln[i] = lens[r];
            ql[i] = qual_lens ? qual_lens[r] : lens[r];
        }
        g = or_bam_pack_fields(n, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], nb, noff, sq,
                               stride, ln, qu, ql, rg, rg_n, perm, is_dup, apply_dup, out, cap);
    }
    free(f); free(noff); free(nb); free(sq); free(qu); free(ln);
    return g;
}

/* ---------------------------------------------------- FASTQ ingest (next row) */

/* Follows _cvt, fq2bam-meth/src/linear_fastq.mojo:37-46: mode 1 (C2T) maps C and
 * c to 'T', mode 2 (G2A) maps G and g to 'A' (upper case out), else as is. */
uint8_t or_fq_cvt(uint8_t c, int mode) {
    if (mode == 1) return (c == 'C' || c == 'c') ? 'T' : c;
    if (mode == 2) return (c == 'G' || c == 'g') ? 'A' : c;
    return c;
}

/* Next line of buf[*pos..n): follows _pipe_line (linear_fastq.mojo:229-266)
 * over a whole in-memory buffer. A line ends at '\n' (one trailing '\r'
 * dropped); at the end, an unterminated rest is a line only when `final` and
 * non-empty after the '\r' strip. Returns 1 with [*s, *e) or 0. */
static int fq_line(const uint8_t *buf, int64_t n, int64_t *pos, int final, int64_t *s, int64_t *e) {
    int64_t i = *pos;
    while (i < n && buf[i] != '\n') i++;
    if (i < n) {
        *s = *pos;
        *e = (i > *pos && buf[i - 1] == '\r') ? i - 1 : i;
        *pos = i + 1;
        return 1;
    }
    if (!final || *pos >= n) return 0;
    *s = *pos;
    *e = (n > *pos && buf[n - 1] == '\r') ? n - 1 : n;
    *pos = n;
    return *e > *s;
}

/* Follows _qname_span (linear_fastq.mojo:322-340): skip a leading '@', stop at
 * the first ' ' or '/'. */
static void fq_qname(const uint8_t *h, int64_t n, int64_t *start, int64_t *len) {
    int64_t s = (n > 0 && h[0] == '@') ? 1 : 0, e = n;
    for (int64_t i = s; i < e; i++)
        if (h[i] == ' ' || h[i] == '/') { e = i; break; }
    *start = s;
    *len = e - s < 0 ? 0 : e - s;
}

/* Parses up to max_records FASTQ records of buf[0..n) (follows _store_rec +
 * fq_read_batch, :342-414, and fq_pack_bases, :432-470): rows[r * stride] =
 * the original bases padded with 'N', conv_rows the same converted with mode
 * (NULL: not written), qual_rows the qualities (first `stride` bytes, 0
 * padded; NULL: not written), lens / qual_lens, names back to back at name_off.
 * info: [0] records, [1] bytes consumed (through the last record's quality
 * line), [2] longest read, [3] status (0 ok; 1, 2, 3 = truncated record
 * missing SEQ, '+', QUAL — only when `final`; 4 = a read longer than stride;
 * 5 = names past names_cap), [4] name bytes. Without `final` an unterminated
 * or incomplete tail is left unconsumed for the next buffer. */
void or_fastq_parse(const uint8_t *buf, int64_t n, int final, int64_t max_records, int64_t stride, int mode,
                    uint8_t *rows, uint8_t *conv_rows, uint8_t *qual_rows, uint32_t *lens,
                    uint32_t *qual_lens, uint8_t *names, int64_t names_cap, int64_t *name_off,
                    int64_t *info) {
    int64_t pos = 0, r = 0, nb = 0, mx = 0, consumed = 0, status = 0;
    name_off[0] = 0;
    while (r < max_records) {
        int64_t hs, he, ss, se, ps, pe, qs, qe;
        if (!fq_line(buf, n, &pos, final, &hs, &he)) break;
        if (!fq_line(buf, n, &pos, final, &ss, &se)) { if (final) status = 1; break; }
        if (!fq_line(buf, n, &pos, final, &ps, &pe)) { if (final) status = 2; break; }
        if (!fq_line(buf, n, &pos, final, &qs, &qe)) { if (final) status = 3; break; }
        (void)ps; (void)pe;
        int64_t slen = se - ss, qlen = qe - qs, q0, ql;
        if (slen > stride) { status = 4; break; }
        fq_qname(buf + hs, he - hs, &q0, &ql);
        if (nb + ql > names_cap) { status = 5; break; }
        memcpy(names + nb, buf + hs + q0, (size_t)ql);
        nb += ql;
        name_off[r + 1] = nb;
        for (int64_t x = 0; x < stride; x++) {
            rows[r * stride + x] = x < slen ? buf[ss + x] : 'N';
            if (conv_rows) conv_rows[r * stride + x] = x < slen ? or_fq_cvt(buf[ss + x], mode) : 'N';
            if (qual_rows) qual_rows[r * stride + x] = x < qlen ? buf[qs + x] : 0;
        }
        lens[r] = (uint32_t)slen;
        qual_lens[r] = (uint32_t)qlen;
        if (slen > mx) mx = slen;
        r++;
        consumed = pos;
    }
    info[0] = r; info[1] = consumed; info[2] = mx; info[3] = status; info[4] = nb;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------ A13: the host (-device cpu) loop */

/* The reference's `-device cpu` mapping of SE reads or of one mate at a time
 * against a dense pack (fq2bam-meth/src/linear_mapper.mojo:209-311, forward
 * strand only, no RC pass):
 *  - the read is converted first (_bs_convert, linear_mapper.mojo:44-61: for
 *    C2T a C or c becomes T, for G2A a G or g becomes A, other bytes are kept);
 *  - every k-mer position i of it is a seed (extract_kmers,
 *    gpu-common/src/gpu_kmer.mojo:6-15), voted as if it sat at query offset
 *    i * seed_stride (vote_dense_seeds, linear_index.mojo:530-578): a k-mer
 *    with a byte outside ACGT/acgt (_encode_kmer_u64, :59-79), absent from the
 *    pack or with more than max_occ postings casts nothing; each posting with
 *    pos >= q_off votes for key (cid, pos - q_off); the winner is the first key
 *    whose count exceeds every count reached before it;
 *  - the winner's window (extend_read_with_seeds, linear_extend.mojo:143-211)
 *    of the read's length from the pack's bytes (none past the contig's end,
 *    contig_window_id, linear_index.mojo:383-394) is compared byte by byte
 *    after upper-casing both; over max_diff mismatches the leading and the
 *    trailing mismatch runs (max_soft each at most) are clipped and the rest
 *    recounted; the hit stands when at most max_diff mismatches remain over at
 *    least 32 aligned bases: flag 0, pos = start + left clip, MAPQ 60 / 40 / 20
 *    for >= 5 / >= 3 / fewer votes.
 * ref_bytes are the pack's sequence bytes (sequences.bin), in the contig_off
 * coordinates of ix. Rows of `stride` bytes; read r uses bs_r1 when r < n_r1,
 * else bs_r2. Pinned by hand-derived reads and by a vote written as the
 * earliest max-count-th occurrence (tests/test_forward.py). */
static inline uint8_t fw_upper(uint8_t c) { return (c >= 'a' && c <= 'z') ? (uint8_t)(c - 32) : c; }

static int64_t fw_mismatches(const uint8_t *w, const uint8_t *q, int64_t from, int64_t to) {
    int64_t nm = 0;
    for (int64_t j = from; j < to; j++) nm += fw_upper(w[j]) != q[j];
    return nm;
}

int or_map_forward(const or_index *ix, const uint8_t *ref_bytes, const uint8_t *reads, int64_t stride,
                   const uint32_t *lens, int64_t n, int64_t n_r1, int bs_r1, int bs_r2, int k, int seed_stride,
                   int64_t max_occ, int max_diff, int max_soft, or_recs *out, int nthreads) {
    if (k < 1 || k > 31 || seed_stride < 1) return -1;
    int status = 0;
#pragma omp parallel num_threads(nthreads)
    {
        uint8_t q[4096];
        int64_t tab_cap = 0;
        uint64_t *tkey = NULL;
        int32_t *tcnt = NULL;
#pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < n; r++) {
            rec_reset(out, r);
            int64_t L = (int64_t)lens[r];
            if (L > stride || L > 4096) L = 0;
            const int mode = r < n_r1 ? bs_r1 : bs_r2;
            const uint8_t *src = reads + r * stride;
            for (int64_t j = 0; j < L; j++) {
                uint8_t u = fw_upper(src[j]);
                if (mode == 1 && u == 'C') u = 'T';
                if (mode == 2 && u == 'G') u = 'A';
                q[j] = u;  /* converted, upper-cased: what the compare sees */
            }
            if (L < k) continue;
            /* postings that can vote: sizes the per-read count table */
            int64_t total = 0;
            for (int64_t i = 0; i + k <= L; i++) {
                uint64_t key = 0;
                int ok = 1;
                for (int j = 0; j < k && ok; j++) {
                    int c = q[i + j] == 'A' ? 0 : q[i + j] == 'C' ? 1 : q[i + j] == 'G' ? 2 : q[i + j] == 'T' ? 3 : -1;
                    if (c < 0) ok = 0;
                    else key = (key << 2) | (uint64_t)c;
                }
                if (!ok) continue;
                uint64_t st, en;
                locate_one(key, ix->kmers, ix->offsets, ix->n_keys, max_occ, &st, &en);
                total += (int64_t)(en - st);
            }
            if (total == 0) continue;
            int64_t cap = 64;
            while (cap < 2 * total) cap <<= 1;
            if (cap > tab_cap) {
                free(tkey);
                free(tcnt);
                tkey = (uint64_t *)malloc((size_t)cap * 8);
                tcnt = (int32_t *)malloc((size_t)cap * 4);
                tab_cap = cap;
                if (!tkey || !tcnt) {
#pragma omp atomic write
                    status = -2;
                    tab_cap = 0;
                    continue;
                }
            }
            memset(tcnt, 0, (size_t)cap * 4);  /* count 0 = empty slot */
            int64_t best_n = 0, best_cid = -1, best_start = 0;
            for (int64_t i = 0; i + k <= L; i++) {
                uint64_t key = 0;
                int ok = 1;
                for (int j = 0; j < k && ok; j++) {
                    int c = q[i + j] == 'A' ? 0 : q[i + j] == 'C' ? 1 : q[i + j] == 'G' ? 2 : q[i + j] == 'T' ? 3 : -1;
                    if (c < 0) ok = 0;
                    else key = (key << 2) | (uint64_t)c;
                }
                if (!ok) continue;
                uint64_t st, en;
                locate_one(key, ix->kmers, ix->offsets, ix->n_keys, max_occ, &st, &en);
                const int64_t q_off = i * seed_stride;
                for (uint64_t p = st; p < en; p++) {
                    const int64_t cid = (int64_t)ix->postings[2 * p];
                    const int64_t pos = (int64_t)ix->postings[2 * p + 1];
                    if (pos < q_off) continue;
                    const uint64_t vk = ((uint64_t)cid << 32) | (uint64_t)(pos - q_off);
                    uint64_t h = (vk * 0x9E3779B97F4A7C15ull) >> 20;
                    while (1) {
                        h &= (uint64_t)(cap - 1);
                        if (tcnt[h] == 0) {
                            tkey[h] = vk;
                            tcnt[h] = 1;
                            break;
                        }
                        if (tkey[h] == vk) {
                            tcnt[h]++;
                            break;
                        }
                        h++;
                    }
                    if (tcnt[h] > best_n) {
                        best_n = tcnt[h];
                        best_cid = cid;
                        best_start = pos - q_off;
                    }
                }
            }
            if (best_cid < 0 || best_n <= 0) continue;
            const int64_t o0 = (int64_t)ix->contig_off[best_cid];
            const int64_t clen = (int64_t)ix->contig_off[best_cid + 1] - o0;
            if (best_start + L > clen) continue;  /* contig_window_id: no window */
            const uint8_t *w = ref_bytes + o0 + best_start;
            int64_t nm = fw_mismatches(w, q, 0, L), sl = 0, sr = 0;
            if (nm > max_diff) {
                while (sl < max_soft && sl < L && fw_upper(w[sl]) != q[sl]) sl++;
                while (sr < max_soft && sr + sl < L && fw_upper(w[L - 1 - sr]) != q[L - 1 - sr]) sr++;
                nm = fw_mismatches(w, q, sl, L - sr);
            }
            const int64_t alen = L - sl - sr;
            if (nm <= max_diff && alen >= 32) {
                out->cid[r] = (int32_t)best_cid;
                out->pos[r] = (uint32_t)(best_start + sl);
                out->mapq[r] = best_n >= 5 ? 60u : best_n >= 3 ? 40u : 20u;
                out->flag[r] = 0;
                out->sl[r] = (uint32_t)sl;
                out->sr[r] = (uint32_t)sr;
                out->nm[r] = (uint32_t)nm;
            }
        }
        free(tkey);
        free(tcnt);
    }
    return status;
}
