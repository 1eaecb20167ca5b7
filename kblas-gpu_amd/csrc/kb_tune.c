/* kb_tune.c -- per-size-bucket kernel choice for the batched Cholesky family.
 *
 * The reference fixes its launch shapes in code (src/batch_triangular/
 * Xpotrf_batch_drivers.cuh:45-61, Xtrsm_batch_drivers.cuh:72-83,
 * Xsyrk_batch_drivers.cuh:162-170) with a compile-time tuning array behind
 * KBLAS_ENABLE_BACKDOORS (src/kblas_struct.h:36-39).  Here the choice is a
 * table of (precision, operation, order range) -> variant, measured on the
 * device by kblas-gpu_amd/tune.py and persisted as JSON; the library loads
 * it once (KBLAS_TUNING_FILE, else tuning.json beside the library) and the
 * dispatch in kb_api.c consults it.  Without a table (or for orders it does
 * not cover) the built-in rule applies.
 *
 * Variants (how many wavefronts serve one matrix and where it lives):
 *   rg16, rg32  four matrices per wavefront, one per 16-lane DPP row, whole
 *               matrix in registers (n <= 16 / n <= 32)
 *   rg4         one matrix per DPP row, 16-column blocks, POTRF, 32 < n <= 64
 *   tile        one matrix per wavefront as ten 16x16 register tiles with
 *               matrix-core updates, n <= 64 (mixed POSV: the tile kernel)
 *   reg         mixed POSV: one matrix per 16/32 lanes, n <= 32
 *   blk         one 256-thread workgroup per matrix, LDS-staged panels, any n
 * Operation "vsplit" (non-uniform batches): the variant is the order at and
 * below which a matrix goes to the wavefront phase (0..64); above it the
 * workgroup phase.
 *
 * The table is process-wide; load or change it before concurrent use.
 */
#define _GNU_SOURCE 1   /* dladdr */
#include <ctype.h>
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kb_internal.h"

#define KB_TUNE_MAX 512

typedef struct {
    char prec;
    int op;
    int n_lo, n_hi;
    int variant;
} kb_tune_entry;

static kb_tune_entry g_tab[KB_TUNE_MAX];
static int g_count;
static int g_init;

static const char* const k_ops[] = {"potrf", "posv", "potrs", "sposv", "vsplit"};
static const char* const k_vars[] = {"auto", "rg16", "rg32", "rg4", "tile", "reg", "blk"};

int kb_tune_op_id(const char* op)
{
    for (int i = 0; i < (int)(sizeof(k_ops) / sizeof(k_ops[0])); ++i)
        if (op && strcmp(op, k_ops[i]) == 0) return i;
    return -1;
}

static int var_id(const char* v)
{
    for (int i = 0; i < (int)(sizeof(k_vars) / sizeof(k_vars[0])); ++i)
        if (v && strcmp(v, k_vars[i]) == 0) return i;
    return -1;
}

/* Is the variant able to serve order n for this operation? */
int kb_tune_legal(int op, int n, int variant)
{
    if (op == KB_TUNE_VSPLIT) return variant >= 0 && variant <= KB_REG_MAX_N;
    switch (variant) {
    case KB_VAR_AUTO: return 1;
    case KB_VAR_RG16: return op != KB_TUNE_SPOSV && n >= 1 && n <= 16;
    case KB_VAR_RG32: return op != KB_TUNE_SPOSV && n >= 1 && n <= 32;
    case KB_VAR_RG4: return op == KB_TUNE_POTRF && n > 32 && n <= 64;
    case KB_VAR_TILE: return n >= 1 && n <= 64;
    case KB_VAR_REG: return op == KB_TUNE_SPOSV && n >= 1 && n <= 32;
    case KB_VAR_BLK: return op != KB_TUNE_SPOSV && n >= 1;
    default: return 0;
    }
}

static int tune_add(char prec, int op, int n_lo, int n_hi, int variant)
{
    if ((prec != 'd' && prec != 's') || op < 0 || n_lo > n_hi) return KBLAS_Error_WrongInput;
    if (op != KB_TUNE_VSPLIT && (variant < 0 || variant > KB_VAR_BLK)) return KBLAS_Error_WrongInput;
    if (op != KB_TUNE_VSPLIT && variant != KB_VAR_AUTO)
        for (int n = n_lo; n <= n_hi; ++n)
            if (!kb_tune_legal(op, n, variant)) return KBLAS_Error_WrongInput;
    if (op == KB_TUNE_VSPLIT && !kb_tune_legal(op, 0, variant)) return KBLAS_Error_WrongInput;
    /* a later entry overrides earlier ones on its range: prepend-by-search order */
    if (g_count >= KB_TUNE_MAX) return KBLAS_Error_WrongInput;
    g_tab[g_count].prec = prec;
    g_tab[g_count].op = op;
    g_tab[g_count].n_lo = n_lo;
    g_tab[g_count].n_hi = n_hi;
    g_tab[g_count].variant = variant;
    ++g_count;
    return KBLAS_Success;
}

/* ------------------------------------------------------------------------
 * A small JSON reader: enough for {"entries": [{"prec": "d", "op": "posv",
 * "n_lo": 1, "n_hi": 16, "variant": "rg16", ...}, ...], ...}; other keys and
 * nested values are skipped.
 * ---------------------------------------------------------------------- */
typedef struct {
    const char* s;
    const char* e;
} jcur;

static void jws(jcur* c)
{
    while (c->s < c->e && isspace((unsigned char)*c->s)) ++c->s;
}

static int jstr(jcur* c, char* out, int cap)
{
    jws(c);
    if (c->s >= c->e || *c->s != '"') return 0;
    ++c->s;
    int k = 0;
    while (c->s < c->e && *c->s != '"') {
        char ch = *c->s++;
        if (ch == '\\' && c->s < c->e) ch = *c->s++;
        if (out && k < cap - 1) out[k++] = ch;
    }
    if (c->s >= c->e) return 0;
    ++c->s;
    if (out) out[k] = 0;
    return 1;
}

static int jskip(jcur* c);

static int jskip_container(jcur* c, char open, char close)
{
    ++c->s;
    jws(c);
    if (c->s < c->e && *c->s == close) { ++c->s; return 1; }
    for (;;) {
        if (open == '{') {
            if (!jstr(c, NULL, 0)) return 0;
            jws(c);
            if (c->s >= c->e || *c->s != ':') return 0;
            ++c->s;
        }
        if (!jskip(c)) return 0;
        jws(c);
        if (c->s >= c->e) return 0;
        if (*c->s == ',') { ++c->s; continue; }
        if (*c->s == close) { ++c->s; return 1; }
        return 0;
    }
}

static int jskip(jcur* c)
{
    jws(c);
    if (c->s >= c->e) return 0;
    char ch = *c->s;
    if (ch == '"') return jstr(c, NULL, 0);
    if (ch == '{') return jskip_container(c, '{', '}');
    if (ch == '[') return jskip_container(c, '[', ']');
    while (c->s < c->e && *c->s != ',' && *c->s != '}' && *c->s != ']' && !isspace((unsigned char)*c->s)) ++c->s;
    return 1;
}

static int jnum(jcur* c, long* v)
{
    jws(c);
    char* end = NULL;
    long x = strtol(c->s, &end, 10);
    if (end == c->s) return 0;
    c->s = end;
    *v = x;
    return 1;
}

/* one entry object */
static int jentry(jcur* c)
{
    char prec[8] = "", op[16] = "", var[16] = "", key[32];
    long n_lo = -1, n_hi = -1, split = -1;
    jws(c);
    if (c->s >= c->e || *c->s != '{') return 0;
    ++c->s;
    for (;;) {
        jws(c);
        if (c->s < c->e && *c->s == '}') { ++c->s; break; }
        if (!jstr(c, key, sizeof key)) return 0;
        jws(c);
        if (c->s >= c->e || *c->s != ':') return 0;
        ++c->s;
        int ok;
        if (!strcmp(key, "prec")) ok = jstr(c, prec, sizeof prec);
        else if (!strcmp(key, "op")) ok = jstr(c, op, sizeof op);
        else if (!strcmp(key, "variant")) {
            jws(c);
            if (c->s < c->e && *c->s == '"') ok = jstr(c, var, sizeof var);
            else ok = jnum(c, &split);
        } else if (!strcmp(key, "n_lo")) ok = jnum(c, &n_lo);
        else if (!strcmp(key, "n_hi")) ok = jnum(c, &n_hi);
        else ok = jskip(c);
        if (!ok) return 0;
        jws(c);
        if (c->s < c->e && *c->s == ',') ++c->s;
    }
    int o = kb_tune_op_id(op);
    if (o < 0 || n_lo < 0 || n_hi < n_lo) return 0;
    int v = (o == KB_TUNE_VSPLIT) ? (int)split : var_id(var);
    return tune_add(prec[0], o, (int)n_lo, (int)n_hi, v) == KBLAS_Success;
}

static int parse_table(const char* txt, long len)
{
    const char* p = strstr(txt, "\"entries\"");
    if (!p) return 0;
    jcur c = {p + 9, txt + len};
    jws(&c);
    if (c.s >= c.e || *c.s != ':') return 0;
    ++c.s;
    jws(&c);
    if (c.s >= c.e || *c.s != '[') return 0;
    ++c.s;
    for (;;) {
        jws(&c);
        if (c.s < c.e && *c.s == ']') return 1;
        if (!jentry(&c)) return 0;
        jws(&c);
        if (c.s < c.e && *c.s == ',') ++c.s;
    }
}

int kblas_tuning_load(const char* path)
{
    if (!path) return KBLAS_Error_WrongInput;
    FILE* f = fopen(path, "rb");
    if (!f) return KBLAS_Error_WrongInput;
    fseek(f, 0, SEEK_END);
    long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (len <= 0 || len > (1L << 22)) { fclose(f); return KBLAS_Error_WrongInput; }
    char* buf = (char*)malloc((size_t)len + 1);
    if (!buf) { fclose(f); return KBLAS_Error_Allocation; }
    long got = (long)fread(buf, 1, (size_t)len, f);
    fclose(f);
    buf[got] = 0;
    g_count = 0;
    int ok = parse_table(buf, got);
    free(buf);
    g_init = 1;
    if (!ok) {
        g_count = 0;   /* a malformed table leaves the built-in rule */
        return KBLAS_Error_WrongInput;
    }
    return KBLAS_Success;
}

void kblas_tuning_clear(void)
{
    g_count = 0;
    g_init = 1;
}

int kblas_tuning_set(char prec, const char* op, int n_lo, int n_hi, const char* variant)
{
    g_init = 1;
    int o = kb_tune_op_id(op);
    if (o < 0) return KBLAS_Error_WrongInput;
    int v = (o == KB_TUNE_VSPLIT) ? (variant ? atoi(variant) : -1) : var_id(variant);
    return tune_add(prec, o, n_lo, n_hi, v);
}

/* Load the default table once: $KBLAS_TUNING_FILE, else tuning.json in the
 * directory of this library. */
void kb_tune_init(void)
{
    if (g_init) return;
    g_init = 1;
    const char* env = getenv("KBLAS_TUNING_FILE");
    if (env && *env) {
        (void)kblas_tuning_load(env);
        return;
    }
    Dl_info di;
    if (dladdr((void*)&kb_tune_init, &di) && di.dli_fname) {
        char path[4096];
        const char* slash = strrchr(di.dli_fname, '/');
        size_t dl = slash ? (size_t)(slash - di.dli_fname) : 0;
        if (dl + 16 < sizeof path) {
            memcpy(path, di.dli_fname, dl);
            strcpy(path + dl, dl ? "/tuning.json" : "tuning.json");
            FILE* f = fopen(path, "rb");
            if (f) {
                fclose(f);
                (void)kblas_tuning_load(path);
            }
        }
    }
}

/* Variant for (prec, op, n): the last matching entry wins; KB_VAR_AUTO when
 * none matches (for vsplit: -1). */
int kb_tune_variant(char prec, int op, int n)
{
    kb_tune_init();
    for (int i = g_count - 1; i >= 0; --i) {
        const kb_tune_entry* t = &g_tab[i];
        if (t->prec == prec && t->op == op && n >= t->n_lo && n <= t->n_hi) return t->variant;
    }
    return (op == KB_TUNE_VSPLIT) ? -1 : KB_VAR_AUTO;
}

int kblas_tuning_query(char prec, const char* op, int n)
{
    int o = kb_tune_op_id(op);
    if (o < 0) return KBLAS_Error_WrongInput;
    return kb_tune_variant(prec, o, n);
}

const char* kblas_tuning_variant_name(int variant)
{
    if (variant < 0 || variant >= (int)(sizeof(k_vars) / sizeof(k_vars[0]))) return "";
    return k_vars[variant];
}
