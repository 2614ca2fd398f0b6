/*
 * Oracle: CPU restatement of one sandengine grid-update step.
 *
 * TEST INFRASTRUCTURE ONLY.  The product never links or calls this file; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, as
 * the checker.
 *
 * It restates the reference compute kernel cell for cell:
 *   main()               shaders/compute/falling_sand.glsl:122-183
 *   simulate()           shaders/compute/falling_sand.glsl:61-118
 *   getMargolusOffset    shaders/compute/operations.glsl:25-34
 *   getCell/outOfBounds  shaders/compute/operations.glsl:7-9, 45-57 (SCREEN_IS_BORDER)
 *   guarded swap(Cell)   shaders/compute/operations.glsl:16-23
 *   hash43 / hash4i      shaders/compute/math.glsl:28-38, 74-79
 *   getMaterialFromID    shaders/compute/gen/materials.glsl:79-86
 *   rule_* / apply*Rules shaders/compute/gen/rules.glsl (semantics of
 *                        sandengine-lang/src/parser/rules.rs:47-97)
 * The rules arrive as a small bytecode built by oracle/oracle.py from the
 * generated GLSL condition/action text; this file interprets it.
 *
 * Two evaluation orders are provided: one block per evaluation (writing its
 * in-bounds cells) and the reference's one-invocation-per-cell order, where
 * each cell recomputes its whole block and keeps its own Margolus index.
 *
 * Floating point: rand is (float)h / 4294967296.0f (u32->f32 round to
 * nearest even, exact scale); densities and literals are f32; comparisons are
 * f32.  The circle brush is sqrt(dx*dx + dy*dy) <= r in f32 (pow(x,2.0) is
 * x*x after Mesa's algebraic lowering); built with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAT_EMPTY 0
#define MAT_NULL 1
#define MAT_WALL 2
#define TYPE_NULL 1
#define TYPE_WALL 2

enum {
    OP_CONST = 1, OP_DENS, OP_MATID, OP_TYPEID, OP_ISTYPE, OP_MATEQ_CC, OP_MATEQ_CK,
    OP_RAND, OP_LT, OP_LE, OP_GT, OP_GE, OP_EQ, OP_NE, OP_AND, OP_OR, OP_NOT, OP_NEG
};
enum { ACT_SWAP = 0, ACT_SET = 1 };

#define MAX_MAT 256
#define MAX_TYPES 256
#define MAX_RULES 64
#define MAX_LEVELS 16
#define MAX_ACTS 16
#define MAX_CODE 512

typedef struct { int len; int32_t code[MAX_CODE]; } expr_t;
typedef struct { int n; int32_t op[MAX_ACTS][3]; } acts_t;
typedef struct {
    int has_prob; float prob; expr_t cond; acts_t acts;
} level_t;
typedef struct {
    int kind;               /* 0 mirrored, 1 left, 2 right */
    int has_pre; expr_t pre;
    int nlev; level_t lev[MAX_LEVELS];
    int has_final; acts_t final_acts;
} rule_t;
typedef struct {
    int nmat; int mat_type[MAX_MAT]; float mat_dens[MAX_MAT];
    int ntypes; int tmembers[MAX_TYPES][MAX_TYPES + 1]; /* [0] = count */
    int nrules; rule_t rules[MAX_RULES];
} ruleset_t;

/* ---------------------------------------------------------------- blob decoding */
static int rd_expr(const int32_t *b, int *p, int n, expr_t *e) {
    if (*p >= n) return -1;
    e->len = b[(*p)++];
    if (e->len < 0 || e->len > MAX_CODE || *p + e->len > n) return -1;
    memcpy(e->code, b + *p, sizeof(int32_t) * e->len);
    *p += e->len;
    return 0;
}
static int rd_acts(const int32_t *b, int *p, int n, acts_t *a) {
    if (*p >= n) return -1;
    a->n = b[(*p)++];
    if (a->n < 0 || a->n > MAX_ACTS || *p + 3 * a->n > n) return -1;
    for (int i = 0; i < a->n; i++)
        for (int k = 0; k < 3; k++) a->op[i][k] = b[(*p)++];
    return 0;
}
static float bits_f(int32_t v) { float f; memcpy(&f, &v, 4); return f; }

static int decode(const int32_t *b, int n, ruleset_t *rs) {
    int p = 0;
    memset(rs, 0, sizeof(*rs));
    rs->nmat = b[p++];
    if (rs->nmat < 3 || rs->nmat > MAX_MAT) return -1;
    for (int i = 0; i < rs->nmat; i++) { rs->mat_type[i] = b[p++]; rs->mat_dens[i] = bits_f(b[p++]); }
    rs->ntypes = b[p++];
    if (rs->ntypes < 3 || rs->ntypes > MAX_TYPES) return -1;
    for (int t = 0; t < rs->ntypes; t++) {
        int k = b[p++];
        if (k < 0 || k > MAX_TYPES) return -1;
        rs->tmembers[t][0] = k;
        for (int j = 0; j < k; j++) rs->tmembers[t][1 + j] = b[p++];
    }
    rs->nrules = b[p++];
    if (rs->nrules < 0 || rs->nrules > MAX_RULES) return -1;
    for (int r = 0; r < rs->nrules; r++) {
        rule_t *R = &rs->rules[r];
        R->kind = b[p++];
        R->has_pre = b[p++];
        if (rd_expr(b, &p, n, &R->pre)) return -1;
        R->nlev = b[p++];
        if (R->nlev < 0 || R->nlev > MAX_LEVELS) return -1;
        for (int l = 0; l < R->nlev; l++) {
            level_t *L = &R->lev[l];
            L->has_prob = b[p++];
            L->prob = bits_f(b[p++]);
            if (rd_expr(b, &p, n, &L->cond)) return -1;
            if (rd_acts(b, &p, n, &L->acts)) return -1;
        }
        R->has_final = b[p++];
        if (rd_acts(b, &p, n, &R->final_acts)) return -1;
    }
    return p == n ? 0 : -1;
}

/* ---------------------------------------------------------------- hash43 */
static uint32_t lowbias(uint32_t h) {
    h ^= h >> 16; h *= 0x7feb352dU; h ^= h >> 15; h *= 0x846ca68bU; h ^= h >> 16;
    return h;
}
void se_ref_hash43(uint32_t px, uint32_t py, uint32_t pz, uint32_t out_u[4], float out_f[4]) {
    uint32_t x = px * 461U + py * 2131U + pz * 2131U * 2131U;
    const uint32_t m[4] = {213U, 2131U, 21313U, 213132U};
    for (int k = 0; k < 4; k++) {
        uint32_t h = lowbias(x * m[k]);
        if (out_u) out_u[k] = h;
        if (out_f) out_f[k] = (float)h / 4294967296.0f;
    }
}

/* ---------------------------------------------------------------- interpreter */
static float eval(const ruleset_t *rs, const expr_t *e, const int c[4], const float rnd[4]) {
    float st[64]; int sp = 0;
    for (int i = 0; i < e->len;) {
        int op = e->code[i++];
        float a, b;
        switch (op) {
        case OP_CONST: st[sp++] = bits_f(e->code[i++]); break;
        case OP_DENS: st[sp++] = rs->mat_dens[c[e->code[i++]]]; break;
        case OP_MATID: st[sp++] = (float)c[e->code[i++]]; break;
        case OP_TYPEID: st[sp++] = (float)rs->mat_type[c[e->code[i++]]]; break;
        case OP_ISTYPE: {
            int cell = e->code[i++], t = e->code[i++];
            int ty = rs->mat_type[c[cell]], hit = 0;
            for (int j = 0; j < rs->tmembers[t][0]; j++) hit |= (rs->tmembers[t][1 + j] == ty);
            st[sp++] = hit ? 1.0f : 0.0f;
            break;
        }
        case OP_MATEQ_CC: { int x = e->code[i++], y = e->code[i++]; st[sp++] = c[x] == c[y]; break; }
        case OP_MATEQ_CK: { int x = e->code[i++], k = e->code[i++]; st[sp++] = c[x] == k; break; }
        case OP_RAND: st[sp++] = rnd[e->code[i++]]; break;
        case OP_NOT: st[sp - 1] = (st[sp - 1] == 0.0f) ? 1.0f : 0.0f; break;
        case OP_NEG: st[sp - 1] = -st[sp - 1]; break;
        default:
            b = st[--sp]; a = st[--sp];
            switch (op) {
            case OP_LT: st[sp++] = a < b; break;
            case OP_LE: st[sp++] = a <= b; break;
            case OP_GT: st[sp++] = a > b; break;
            case OP_GE: st[sp++] = a >= b; break;
            case OP_EQ: st[sp++] = a == b; break;
            case OP_NE: st[sp++] = a != b; break;
            case OP_AND: st[sp++] = (a != 0.0f) && (b != 0.0f); break;
            case OP_OR: st[sp++] = (a != 0.0f) || (b != 0.0f); break;
            default: return 0.0f;
            }
        }
    }
    return sp ? st[sp - 1] : 0.0f;
}

static void gswap(const ruleset_t *rs, int *a, int *b) {
    int ta = rs->mat_type[*a], tb = rs->mat_type[*b];
    if (ta == TYPE_WALL || tb == TYPE_WALL || ta == TYPE_NULL || tb == TYPE_NULL) return;
    int t = *a; *a = *b; *b = t;
}

static void run_acts(const ruleset_t *rs, const acts_t *A, int c[4]) {
    for (int i = 0; i < A->n; i++) {
        if (A->op[i][0] == ACT_SWAP) gswap(rs, &c[A->op[i][1]], &c[A->op[i][2]]);
        else c[A->op[i][1]] = A->op[i][2];
    }
}

static void run_rule(const ruleset_t *rs, const rule_t *R, int c[4], const float rnd[4]) {
    if (R->has_pre && eval(rs, &R->pre, c, rnd) == 0.0f) return;
    for (int l = 0; l < R->nlev; l++) {
        const level_t *L = &R->lev[l];
        int ok = 1;
        if (L->has_prob) ok = rnd[1] <= L->prob;
        if (ok) ok = eval(rs, &L->cond, c, rnd) != 0.0f;
        if (ok) { run_acts(rs, &L->acts, c); return; }
    }
    if (R->has_final) run_acts(rs, &R->final_acts, c);
}

static int get_cell(const ruleset_t *rs, const uint32_t *g, int W, int H, int x, int y) {
    if (x < 0 || y < 0 || x >= W || y >= H) return MAT_WALL;
    uint32_t v = g[(size_t)y * W + x];
    return v < (uint32_t)rs->nmat ? (int)v : MAT_NULL;
}

/* simulate() after the block's cells are read (falling_sand.glsl:72-118): the
 * all-EMPTY early-out, then mirror, rules and un-mirror with the given rand. */
static void sim_cells(const ruleset_t *rs, int c[4], const float rnd[4]) {
    if (c[0] == MAT_EMPTY && c[1] == MAT_EMPTY && c[2] == MAT_EMPTY && c[3] == MAT_EMPTY) return;
    int mirror = rnd[0] < 0.5f;
    if (mirror) { gswap(rs, &c[0], &c[1]); gswap(rs, &c[2], &c[3]); }
    for (int r = 0; r < rs->nrules; r++)
        if (rs->rules[r].kind == 0) run_rule(rs, &rs->rules[r], c, rnd);
    if (mirror) { gswap(rs, &c[0], &c[1]); gswap(rs, &c[2], &c[3]); }
    for (int r = 0; r < rs->nrules; r++) {
        int k = rs->rules[r].kind;
        if ((!mirror && k == 2) || (mirror && k == 1)) run_rule(rs, &rs->rules[r], c, rnd);
    }
}

/* simulate() for the block with origin (ox,oy): fills c[0..3] with the result. */
static void sim_block(const ruleset_t *rs, const uint32_t *g, int W, int H, int ox, int oy,
                      int frame, int c[4]) {
    c[0] = get_cell(rs, g, W, H, ox, oy);
    c[1] = get_cell(rs, g, W, H, ox + 1, oy);
    c[2] = get_cell(rs, g, W, H, ox, oy + 1);
    c[3] = get_cell(rs, g, W, H, ox + 1, oy + 1);
    float rnd[4];
    se_ref_hash43((uint32_t)ox, (uint32_t)oy, (uint32_t)frame, NULL, rnd);
    sim_cells(rs, c, rnd);
}


/* falling_sand.glsl:134-179: returns the override material or -1. */
static int mod_override(const ruleset_t *rs, const int32_t *mods, int nmods, int x, int y) {
    int got = 0, fin = MAT_NULL;
    for (int i = 0; i < nmods && i < 256; i++) {
        const int32_t *m = mods + 8 * i;
        if (m[3] == 0) break;
        int dx = abs(m[0] - x), dy = abs(m[1] - y);
        int mat = (m[4] >= 0 && m[4] < rs->nmat) ? m[4] : MAT_NULL;
        if (m[2] == 0) {
            float fx = (float)dx, fy = (float)dy;
            float dist = sqrtf(fx * fx + fy * fy);
            if (dist <= (float)m[3]) { got = 1; fin = mat; }
        } else if (m[2] == 1) {
            if (dx <= m[3] && dy <= m[3]) { got = 1; fin = mat; }
        }
    }
    return (got && fin != MAT_NULL) ? fin : -1;
}

static void offset(int frame, int *offx, int *offy) {
    int f = frame % 4;
    if (f == 1) { *offx = 1; *offy = 1; }
    else if (f == 2) { *offx = 0; *offy = 1; }
    else if (f == 3) { *offx = 1; *offy = 0; }
    else { *offx = 0; *offy = 0; }
}

static ruleset_t *g_cache_rs = NULL;
static int32_t *g_cache_blob = NULL;
static int g_cache_len = -1;

static const ruleset_t *get_rs(const int32_t *blob, int len) {
    if (g_cache_rs && g_cache_len == len && memcmp(g_cache_blob, blob, 4 * (size_t)len) == 0)
        return g_cache_rs;
    ruleset_t *rs = (ruleset_t *)malloc(sizeof(ruleset_t));
    if (!rs || decode(blob, len, rs)) { free(rs); return NULL; }
    free(g_cache_rs); free(g_cache_blob);
    g_cache_blob = (int32_t *)malloc(4 * (size_t)len);
    memcpy(g_cache_blob, blob, 4 * (size_t)len);
    g_cache_len = len;
    g_cache_rs = rs;
    return rs;
}

int se_ref_validate(const int32_t *blob, int len) { return get_rs(blob, len) ? 0 : -1; }

/* One step: `frame` is the value the reference's Simulation::run dispatches with
 * (already incremented).  percell=1 evaluates one block per cell, as the
 * reference's invocations do. */
int se_ref_step(const int32_t *blob, int len, const uint32_t *gin, uint32_t *gout, int W, int H,
                int frame, const int32_t *mods, int nmods, int percell) {
    const ruleset_t *rs = get_rs(blob, len);
    if (!rs || W <= 0 || H <= 0) return -1;
    if (frame == 1) { memset(gout, 0, sizeof(uint32_t) * (size_t)W * H); return 0; }
    int offx, offy;
    offset(frame, &offx, &offy);
    if (percell) {
        #pragma omp parallel for schedule(static)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                int ov = nmods ? mod_override(rs, mods, nmods, x, y) : -1;
                if (ov >= 0) { gout[(size_t)y * W + x] = (uint32_t)ov; continue; }
                int qx = x + offx, qy = y + offy;
                int ox = (qx / 2) * 2 - offx, oy = (qy / 2) * 2 - offy;
                int k = (qx & 1) + 2 * (qy & 1);
                int c[4];
                sim_block(rs, gin, W, H, ox, oy, frame, c);
                gout[(size_t)y * W + x] = (uint32_t)c[k];
            }
        return 0;
    }
    int nby = (H + offy + 1) / 2, nbx = (W + offx + 1) / 2;
    #pragma omp parallel for schedule(static)
    for (int by = 0; by < nby; by++) {
        int oy = 2 * by - offy;
        for (int bx = 0; bx < nbx; bx++) {
            int ox = 2 * bx - offx;
            int c[4];
            sim_block(rs, gin, W, H, ox, oy, frame, c);
            for (int k = 0; k < 4; k++) {
                int x = ox + (k & 1), y = oy + (k >> 1);
                if (x < 0 || y < 0 || x >= W || y >= H) continue;
                int ov = nmods ? mod_override(rs, mods, nmods, x, y) : -1;
                gout[(size_t)y * W + x] = (uint32_t)(ov >= 0 ? ov : c[k]);
            }
        }
    }
    return 0;
}

/* Runs frames first_frame .. first_frame+nsteps-1 with ping-pong buffers;
 * mods apply to the first of them only.  Result lands in `grid`. */
int se_ref_run(const int32_t *blob, int len, uint32_t *grid, int W, int H, int first_frame,
               int nsteps, const int32_t *mods, int nmods) {
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)W * H);
    if (!tmp) return -1;
    uint32_t *a = grid, *b = tmp;
    for (int s = 0; s < nsteps; s++) {
        int rc = se_ref_step(blob, len, a, b, W, H, first_frame + s, s == 0 ? mods : NULL,
                             s == 0 ? nmods : 0, 0);
        if (rc) { free(tmp); return rc; }
        uint32_t *t = a; a = b; b = t;
    }
    if (a != grid) memcpy(grid, a, sizeof(uint32_t) * (size_t)W * H);
    free(tmp);
    return 0;
}

/* Every block of ids in [0, nid) under each of `nh` hash43 outputs (4 words
 * each, rand = (float)h / 2^32 as in math.glsl:74-79): out[((t * nh) + k) * 4 + i]
 * is cell i of the result for tuple t = c0 + nid*c1 + nid^2*c2 + nid^3*c3 under
 * hash k.  For exhaustive comparison with other block evaluators. */
int se_ref_block_table(const int32_t *blob, int len, int nid, const uint32_t *hashes, int nh, uint8_t *out) {
    const ruleset_t *rs = get_rs(blob, len);
    if (!rs || nid < 1 || nid > rs->nmat || nid > 256 || nh < 1) return -1;
    const long nt = (long)nid * nid * nid * nid;
    #pragma omp parallel for schedule(static)
    for (long t = 0; t < nt; t++) {
        for (int k = 0; k < nh; k++) {
            int c[4] = {(int)(t % nid), (int)(t / nid % nid), (int)(t / nid / nid % nid), (int)(t / nid / nid / nid)};
            float rnd[4];
            for (int l = 0; l < 4; l++) rnd[l] = (float)hashes[4 * k + l] / 4294967296.0f;
            sim_cells(rs, c, rnd);
            for (int i = 0; i < 4; i++) out[(t * nh + k) * 4 + i] = (uint8_t)c[i];
        }
    }
    return 0;
}
