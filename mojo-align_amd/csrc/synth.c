/*
 * Seeded synthetic WGBS workloads: references, bisulfite reads and sort
 * records. Counter-based randomness (splitmix64 finaliser over (seed, stream,
 * index)) so every element is independent of thread count and generation
 * order: the same seed gives the same bytes on any machine.
 *
 * Read model (single-end, R1 convention): a template is drawn uniformly from a
 * contig and strand (OT = top strand as is, OB = reverse complement of the top
 * strand); every C of the template in CpG context (next template base G) stays
 * C with p_meth, every other C becomes T with p_conv; then each base is
 * substituted with p_err by one of the other three bases; a dup_frac share of
 * reads are exact copies of an earlier read (planted PCR duplicates).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Random 64-bit word for (seed, stream, index). */
static inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(seed * 0x9E3779B97F4A7C15ull + mix64(stream + 0x632BE59BD9B4E019ull) +
                 i * 0xD1B54A32D192ED03ull);
}

static inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

static const uint8_t BASES[4] = {'A', 'C', 'G', 'T'};

static inline uint8_t comp(uint8_t b) {
    switch (b) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return 'N';
    }
}

/* Contig lengths: equal split of total_len (remainder to the last contig). */
void ms_contig_layout(int64_t n_contigs, int64_t total_len, uint64_t *contig_off) {
    int64_t each = total_len / n_contigs;
    for (int64_t c = 0; c <= n_contigs; c++) contig_off[c] = (uint64_t)(c * each);
    contig_off[n_contigs] = (uint64_t)total_len;
}

/*
 * Reference bases (uppercase ACGT) into out[total_len].
 *   cpg_keep     fraction of random CG dinucleotides kept (others C->T), so the
 *                CpG share of dinucleotides is about cpg_keep/16 (0.16 -> 1%).
 *   repeat_frac  share of bases overwritten by tandem-repeat blocks (unit 2-6 bp,
 *                block 100-1000 bp) to create multi-hit seeds.
 *   segdup_*     interspersed segmental duplicates: `families` source segments
 *                of `seg_len` bp copied `copies` times each with per-base
 *                divergence `div` (creates seeds with occ in the hundreds).
 */
int ms_gen_reference(uint64_t seed, int64_t n_contigs, int64_t total_len, double cpg_keep,
                     double repeat_frac, int64_t seg_families, int64_t seg_copies,
                     int64_t seg_len, double seg_div, uint8_t *out, uint64_t *contig_off) {
    if (n_contigs < 1 || total_len < n_contigs) return -1;
    ms_contig_layout(n_contigs, total_len, contig_off);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < total_len; i++) out[i] = BASES[rnd(seed, 1, (uint64_t)i) & 3];
    /* tandem repeats: block b starts at a random position, fills len bases */
    if (repeat_frac > 0.0) {
        int64_t target = (int64_t)(repeat_frac * (double)total_len);
        int64_t covered = 0;
        for (uint64_t b = 0; covered < target; b++) {
            uint64_t r0 = rnd(seed, 2, b * 4), r1 = rnd(seed, 2, b * 4 + 1);
            uint64_t r2 = rnd(seed, 2, b * 4 + 2), r3 = rnd(seed, 2, b * 4 + 3);
            int64_t len = 100 + (int64_t)(r0 % 901);
            int64_t start = (int64_t)(r1 % (uint64_t)(total_len - len > 0 ? total_len - len : 1));
            int unit_len = 2 + (int)(r2 % 5);
            uint8_t unit_b[6];
            for (int u = 0; u < unit_len; u++) unit_b[u] = BASES[(r3 >> (2 * u)) & 3];
            for (int64_t j = 0; j < len && start + j < total_len; j++)
                out[start + j] = unit_b[j % unit_len];
            covered += len;
        }
    }
    /* segmental duplicates with divergence */
    if (seg_families > 0 && seg_copies > 0 && seg_len > 0 && seg_len < total_len) {
        for (int64_t f = 0; f < seg_families; f++) {
            int64_t src = (int64_t)(rnd(seed, 3, (uint64_t)f) % (uint64_t)(total_len - seg_len));
            uint8_t *tmpl = (uint8_t *)malloc((size_t)seg_len);
            if (!tmpl) return -2;
            memcpy(tmpl, out + src, (size_t)seg_len);
            for (int64_t c = 0; c < seg_copies; c++) {
                uint64_t key = (uint64_t)(f * seg_copies + c);
                int64_t dst = (int64_t)(rnd(seed, 4, key) % (uint64_t)(total_len - seg_len));
                for (int64_t j = 0; j < seg_len; j++) {
                    uint8_t b = tmpl[j];
                    uint64_t r = rnd(seed, 5 + (uint64_t)f, (uint64_t)(c * seg_len + j));
                    if (unit(r) < seg_div) b = BASES[(r >> 3) & 3];
                    out[dst + j] = b;
                }
            }
            free(tmpl);
        }
    }
    /* CpG depletion: CG -> TG with probability 1 - cpg_keep (creates no new CG) */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < total_len - 1; i++) {
        if (out[i] == 'C' && out[i + 1] == 'G' && unit(rnd(seed, 6, (uint64_t)i)) >= cpg_keep)
            out[i] = 'T';
    }
    return 0;
}

/* bwameth-style dual conversion: for each contig c, f-copy (C->T) at 2c and
 * r-copy (G->A) at 2c+1. out holds 2*total bases, out_off 2*n+1 offsets. */
void ms_dual_convert(const uint8_t *ref, const uint64_t *contig_off, int64_t n_contigs,
                     uint8_t *out, uint64_t *out_off) {
    uint64_t o = 0;
    for (int64_t c = 0; c < n_contigs; c++) {
        uint64_t a = contig_off[c], b = contig_off[c + 1], len = b - a;
        out_off[2 * c] = o;
        out_off[2 * c + 1] = o + len;
        o += 2 * len;
    }
    out_off[2 * n_contigs] = o;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < n_contigs; c++) {
        uint64_t a = contig_off[c], len = contig_off[c + 1] - a;
        uint8_t *f = out + out_off[2 * c], *r = out + out_off[2 * c + 1];
        for (uint64_t i = 0; i < len; i++) {
            uint8_t b = ref[a + i];
            f[i] = (b == 'C' || b == 'c') ? 'T' : b;
            r[i] = (b == 'G' || b == 'g') ? 'A' : b;
        }
    }
}

/*
 * Reads into out[n_reads x max_len] (ASCII, padded 'N'), lens[n_reads].
 * Lengths uniform in [len_min, len_max]. Truth per read: contig (original
 * numbering), 0-based template start on the top strand, strand (0 OT, 1 OB),
 * dup_of (-1 or the earlier read copied).
 */
int ms_gen_reads(uint64_t seed, const uint8_t *ref, const uint64_t *contig_off, int64_t n_contigs,
                 int64_t n_reads, int32_t len_min, int32_t len_max, double p_conv, double p_meth,
                 double p_err, double dup_frac, int64_t max_len, uint8_t *out, uint32_t *lens,
                 int32_t *t_contig, uint32_t *t_pos, uint8_t *t_strand, int32_t *dup_of) {
    if (len_min < 1 || len_max < len_min || max_len < len_max) return -1;
    int64_t total = (int64_t)contig_off[n_contigs];
    for (int64_t c = 0; c < n_contigs; c++)
        if ((int64_t)(contig_off[c + 1] - contig_off[c]) <= len_max + 1) return -2;
    int fail = 0;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n_reads; r++) {
        uint64_t s = (uint64_t)r * 8;
        int32_t L = len_min + (int32_t)(rnd(seed, 10, s) % (uint64_t)(len_max - len_min + 1));
        /* contig by length: global position, then locate its contig */
        int64_t c = 0;
        int64_t g = (int64_t)(rnd(seed, 10, s + 1) % (uint64_t)total);
        while (c + 1 < n_contigs && (int64_t)contig_off[c + 1] <= g) c++;
        int64_t clen = (int64_t)(contig_off[c + 1] - contig_off[c]);
        int64_t start = (int64_t)(rnd(seed, 10, s + 2) % (uint64_t)(clen - L + 1));
        int strand = (int)(rnd(seed, 10, s + 3) & 1);
        const uint8_t *cs = ref + contig_off[c];
        uint8_t *dst = out + r * max_len;
        /* template plus one base of downstream context for the CpG test */
        uint8_t tmpl[1024 + 1];
        if (L > 1024) { fail = 1; continue; }
        for (int32_t j = 0; j <= L; j++) {
            uint8_t b;
            if (strand == 0) {
                int64_t p = start + j;
                b = p < clen ? cs[p] : 'N';
            } else {
                int64_t p = start + L - 1 - j;
                b = p >= 0 ? comp(cs[p]) : 'N';
            }
            tmpl[j] = b;
        }
        for (int32_t j = 0; j < L; j++) {
            uint8_t b = tmpl[j];
            if (b == 'C') {
                double u = unit(rnd(seed, 11, (uint64_t)r * 1024 + (uint64_t)j));
                if (tmpl[j + 1] == 'G') b = (u < p_meth) ? 'C' : 'T';
                else b = (u < p_conv) ? 'T' : 'C';
            }
            uint64_t re = rnd(seed, 12, (uint64_t)r * 1024 + (uint64_t)j);
            if (unit(re) < p_err) {
                int cur = b == 'A' ? 0 : b == 'C' ? 1 : b == 'G' ? 2 : 3;
                b = BASES[(cur + 1 + (int)((re >> 3) % 3)) & 3];
            }
            dst[j] = b;
        }
        for (int64_t j = L; j < max_len; j++) dst[j] = 'N';
        lens[r] = (uint32_t)L;
        t_contig[r] = (int32_t)c;
        t_pos[r] = (uint32_t)start;
        t_strand[r] = (uint8_t)strand;
        dup_of[r] = -1;
    }
    if (fail) return -3;
    /* planted duplicates, in read order so chains resolve to final content */
    for (int64_t r = 1; r < n_reads; r++) {
        uint64_t rd = rnd(seed, 13, (uint64_t)r);
        if (unit(rd) < dup_frac) {
            int64_t j = (int64_t)(rnd(seed, 14, (uint64_t)r) % (uint64_t)r);
            memcpy(out + r * max_len, out + j * max_len, (size_t)max_len);
            lens[r] = lens[j];
            t_contig[r] = t_contig[j];
            t_pos[r] = t_pos[j];
            t_strand[r] = t_strand[j];
            dup_of[r] = (int32_t)j;
        }
    }
    return 0;
}

/* 2-bit pack, MSB-first within each byte (base 4i in bits 7-6), A0 C1 G2 T3.
 * Returns the number of bases that were not ACGT (packed as A). */
int64_t ms_pack_2bit(const uint8_t *ascii, const uint32_t *lens, int64_t n, int64_t max_len,
                     int64_t stride_bytes, uint8_t *out) {
    int64_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t r = 0; r < n; r++) {
        uint8_t *d = out + r * stride_bytes;
        memset(d, 0, (size_t)stride_bytes);
        const uint8_t *s = ascii + r * max_len;
        for (uint32_t j = 0; j < lens[r]; j++) {
            uint8_t b = s[j], c;
            switch (b) {
            case 'A': case 'a': c = 0; break;
            case 'C': case 'c': c = 1; break;
            case 'G': case 'g': c = 2; break;
            case 'T': case 't': c = 3; break;
            default: c = 0; bad++;
            }
            d[j >> 2] |= (uint8_t)(c << (6 - 2 * (j & 3)));
        }
    }
    return bad;
}

/*
 * Sort-stress records: key = contig<<40 | pos<<1 | strand over a genome of
 * `genome_len` bases in n_contigs equal contigs; read id = index; nm in 0..8.
 * A dup_frac share of records sit in clusters of size 2..8 sharing one key.
 */
void ms_gen_sort_records(uint64_t seed, int64_t n, int64_t n_contigs, int64_t genome_len,
                         double dup_frac, uint64_t *key, uint32_t *rid, uint8_t *nm) {
    int64_t clen = genome_len / n_contigs;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint64_t g = rnd(seed, 20, (uint64_t)i) % (uint64_t)genome_len;
        uint64_t c = g / (uint64_t)clen;
        if (c >= (uint64_t)n_contigs) c = (uint64_t)n_contigs - 1;
        uint64_t p = g - c * (uint64_t)clen;
        key[i] = (c << 40) | (p << 1) | (rnd(seed, 21, (uint64_t)i) & 1);
        rid[i] = (uint32_t)i;
        nm[i] = (uint8_t)(rnd(seed, 22, (uint64_t)i) % 9);
    }
    /* clusters: walk records; a cluster head copies its key to the next 1..7 */
    int64_t i = 0;
    while (i < n) {
        uint64_t r = rnd(seed, 23, (uint64_t)i);
        if (unit(r) < dup_frac / 5.0) {
            int64_t sz = 2 + (int64_t)((r >> 7) % 7);
            for (int64_t j = 1; j < sz && i + j < n; j++) {
                int64_t t = i + j * 7919 % (n - i);
                key[t] = key[i];
            }
            i += sz;
        } else {
            i++;
        }
    }
}

/* FASTQ-like read names and qualities for n reads: name i is an
 * Illumina-style "SIM:<run>:HMI355X:<lane>:<tile>:<x>:<y>" (names back to back,
 * name_off[n+1]); qualities are phred+33 bytes into [n x max_len] rows, Q2..Q41
 * with a slow decline along the read and occasional dips. Returns the bytes of
 * names written, or -1 when name_cap is too small. */
int64_t ms_gen_names_quals(uint64_t seed, int64_t n, const uint32_t *lens, int64_t max_len,
                           uint8_t *qual, char *names, int64_t name_cap, int64_t *name_off) {
    int64_t off = 0;
    for (int64_t i = 0; i < n; i++) {
        uint64_t r = rnd(seed, 40, (uint64_t)i);
        int len = 0;
        char buf[64];
        unsigned lane = 1 + (unsigned)(r % 4), tile = 1101 + (unsigned)((r >> 8) % 24);
        unsigned x = (unsigned)((r >> 16) % 30000), y = (unsigned)((r >> 32) % 30000);
        len = snprintf(buf, sizeof buf, "SIM:%u:HMI355X:%u:%u:%u:%u", 1 + (unsigned)(i / 100000000), lane, tile, x, y);
        if (off + len > name_cap) return -1;
        name_off[i] = off;
        memcpy(names + off, buf, (size_t)len);
        off += len;
    }
    name_off[n] = off;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint8_t *q = qual + i * max_len;
        int64_t L = lens[i];
        for (int64_t j = 0; j < max_len; j++) {
            if (j >= L) { q[j] = 0; continue; }
            uint64_t r = rnd(seed, 41 + (uint64_t)i, (uint64_t)j);
            int base = 40 - (int)(j * 8 / (L > 0 ? L : 1));
            int v = base - (int)(r % 6);
            if ((r >> 20) % 50 == 0) v = 2 + (int)((r >> 28) % 12);
            if (v < 2) v = 2;
            if (v > 41) v = 41;
            q[j] = (uint8_t)(33 + v);
        }
    }
    return off;
}

/* FASTQ text of n reads: "@name\nbases\n+\nqualities\n" per read (bases and
 * qualities are rows of stride bytes, lens[i] used). out_off[n+1] receives
 * each record's offset; returns the text length, or -1 past cap. */
int64_t ms_write_fastq(int64_t n, const uint8_t *bases, const uint8_t *qual, int64_t stride,
                       const uint32_t *lens, const char *names, const int64_t *name_off, char *out, int64_t cap,
                       int64_t *out_off) {
    int64_t off = 0;
    for (int64_t i = 0; i < n; i++) {
        out_off[i] = off;
        off += 1 + (name_off[i + 1] - name_off[i]) + 1 + 2 * (int64_t)lens[i] + 4;
    }
    out_off[n] = off;
    if (off > cap) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        char *p = out + out_off[i];
        const int64_t nl = name_off[i + 1] - name_off[i], L = lens[i];
        *p++ = '@';
        memcpy(p, names + name_off[i], (size_t)nl);
        p += nl;
        *p++ = '\n';
        memcpy(p, bases + i * stride, (size_t)L);
        p += L;
        *p++ = '\n';
        *p++ = '+';
        *p++ = '\n';
        memcpy(p, qual + i * stride, (size_t)L);
        p += L;
        *p++ = '\n';
    }
    return off;
}
