/*
 * sk_oracle.c — CPU oracle (TEST INFRASTRUCTURE ONLY; never part of the
 * product path).  A plain-C restatement of the reference interpreter
 * strata-eval: naive T_P (crates/strata-eval/src/naive.rs) and semi-naive
 * delta evaluation (crates/strata-eval/src/seminaive.rs), with the relation
 * store of store.rs and the annotations of value.rs / strata-ir trop.rs.
 *
 * What is restated exactly:
 *   - rules per stratum in declaration order; aggregate rules once, first
 *     (naive.rs:175-181, seminaive.rs:79-84);
 *   - the nested-loop join in body order with a fresh binding copy per row,
 *     ⊗ accumulated left to right from otimes_id(head semiring)
 *     (seminaive.rs:215-298, naive.rs:237-297), every unify of naive.rs:407-455;
 *   - Jacobi application: every derivation of the round is collected first,
 *     then ⊕-combined (seminaive.rs:192-201, naive.rs:220-224);
 *   - next delta = the merged value of each changed tuple (seminaive.rs:203-208);
 *   - round counters and the Bellman-Ford cap universe+2 checked after every
 *     round (seminaive.rs:94-104, naive.rs:185-231);
 *   - every EvalError variant at the point the reference raises it, and the
 *     reference's panics (annotation mismatch, value.rs:56-62) as ORC_PANIC.
 *
 * Orders the reference leaves open: its delta rows come from a HashSet
 * (seminaive.rs:193), randomly ordered per process; here they are visited in
 * sorted order, which is one admissible run of the reference.  Full relations
 * are visited in BTreeMap (sorted) order exactly as the reference does.
 *
 * Not evaluated: compound terms (@terms), which the reference routes to the
 * naive engine with a term table; a compound head or negated atom returns
 * ORC_ERR_UNSUPPORTED.  A body compound never unifies with a Sym/Int value
 * (naive.rs:439-443), which is restated as is.
 */
#define _GNU_SOURCE
#include "sk_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ values */

static inline int val_cmp(const orc_val* a, const orc_val* b) {
    if (a->kind != b->kind) return a->kind < b->kind ? -1 : 1;
    if (a->v != b->v) return a->v < b->v ? -1 : 1;
    return 0;
}

static inline int val_eq(const orc_val* a, const orc_val* b) {
    return a->kind == b->kind && a->v == b->v;
}

static int tup_cmp(const orc_val* a, const orc_val* b, uint32_t arity) {
    for (uint32_t i = 0; i < arity; i++) {
        int c = val_cmp(a + i, b + i);
        if (c) return c;
    }
    return 0;
}

static inline uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}

static uint64_t tup_hash(const orc_val* t, uint32_t arity) {
    uint64_t h = 0x9e3779b97f4a7c15ULL ^ arity;
    for (uint32_t i = 0; i < arity; i++)
        h = mix64(h ^ ((uint64_t)t[i].v * 2654435761ULL) ^ ((uint64_t)t[i].kind << 61));
    return h;
}

/* ------------------------------------------------------------- annotations */

/* Weight order: every finite value below +inf (trop.rs:169-175). */
static inline int w_lt(const orc_ann* a, const orc_ann* b) {
    if (a->inf) return 0;
    if (b->inf) return 1;
    return a->w < b->w;
}

/* Ann::oplus (value.rs:43-52): Bool OR is idempotent and never "changes";
 * Trop keeps the min and reports a change only on a strictly smaller weight. */
static int ann_oplus(orc_ann existing, orc_ann other, orc_ann* merged, int* changed) {
    if (existing.kind == ORC_ANN_UNIT && other.kind == ORC_ANN_UNIT) {
        *merged = existing;
        *changed = 0;
        return ORC_OK;
    }
    if (existing.kind == ORC_ANN_W && other.kind == ORC_ANN_W) {
        orc_ann m = w_lt(&other, &existing) ? other : existing;
        *merged = m;
        *changed = w_lt(&m, &existing);
        return ORC_OK;
    }
    return ORC_PANIC;
}

/* Ann::otimes (value.rs:56-62) over Weight::otimes (trop.rs:56-62): +inf
 * absorbs, finite addition is checked and overflow is an error. */
static int ann_otimes(orc_ann a, orc_ann b, orc_ann* out, orc_error* err) {
    if (a.kind == ORC_ANN_UNIT && b.kind == ORC_ANN_UNIT) {
        *out = a;
        return ORC_OK;
    }
    if (a.kind == ORC_ANN_W && b.kind == ORC_ANN_W) {
        memset(out, 0, sizeof *out);
        out->kind = ORC_ANN_W;
        if (a.inf || b.inf) {
            out->inf = 1;
            out->w = 0;
            return ORC_OK;
        }
        int64_t r;
        if (__builtin_add_overflow(a.w, b.w, &r)) {
            if (err) {
                err->code = ORC_ERR_OVERFLOW;
                err->lhs = a.w;
                err->rhs = b.w;
                snprintf(err->msg, sizeof err->msg,
                         "tropical weight overflow: %lld + %lld exceeds i64",
                         (long long)a.w, (long long)b.w);
            }
            return ORC_ERR_OVERFLOW;
        }
        out->w = r;
        return ORC_OK;
    }
    if (err) {
        err->code = ORC_PANIC;
        snprintf(err->msg, sizeof err->msg, "semiring annotation mismatch in otimes");
    }
    return ORC_PANIC;
}

static orc_ann ann_unit(void) {
    orc_ann a;
    memset(&a, 0, sizeof a);
    a.kind = ORC_ANN_UNIT;
    return a;
}

static orc_ann ann_w(int64_t w) {
    orc_ann a;
    memset(&a, 0, sizeof a);
    a.kind = ORC_ANN_W;
    a.w = w;
    return a;
}

/* Ann::otimes_id (value.rs:33-40): Bool Unit, Trop Finite(0). */
static orc_ann ann_otimes_id(uint32_t sem) {
    return sem == ORC_TROP ? ann_w(0) : ann_unit();
}

/* ------------------------------------------------------------------ rowset */

/* A set of rows with a lazily maintained sorted view (BTreeMap order) and
 * lazily built per-column position indexes over that view. */
typedef struct {
    uint32_t arity;
    uint64_t n, cap;
    orc_val* vals;
    orc_ann* anns;
    uint32_t* order;     /* row ids in sorted order, covers [0, order_n) */
    uint64_t order_n;
    uint32_t** colidx;   /* per column: positions into order, sorted by (value, position) */
    uint64_t* colidx_n;  /* colidx[j] valid iff colidx_n[j] == n == order_n */
} rowset;

static void rs_init(rowset* rs, uint32_t arity) {
    memset(rs, 0, sizeof *rs);
    rs->arity = arity;
    if (arity) {
        rs->colidx = (uint32_t**)calloc(arity, sizeof(uint32_t*));
        rs->colidx_n = (uint64_t*)calloc(arity, sizeof(uint64_t));
    }
}

static void rs_free(rowset* rs) {
    free(rs->vals);
    free(rs->anns);
    free(rs->order);
    if (rs->colidx)
        for (uint32_t j = 0; j < rs->arity; j++) free(rs->colidx[j]);
    free(rs->colidx);
    free(rs->colidx_n);
    memset(rs, 0, sizeof *rs);
}

static int rs_reserve(rowset* rs, uint64_t want) {
    if (want <= rs->cap) return 0;
    uint64_t nc = rs->cap ? rs->cap : 64;
    while (nc < want) nc *= 2;
    orc_val* v = (orc_val*)realloc(rs->vals, nc * (rs->arity ? rs->arity : 1) * sizeof(orc_val));
    if (!v) return -1;
    rs->vals = v;
    orc_ann* a = (orc_ann*)realloc(rs->anns, nc * sizeof(orc_ann));
    if (!a) return -1;
    rs->anns = a;
    uint32_t* o = (uint32_t*)realloc(rs->order, nc * sizeof(uint32_t));
    if (!o) return -1;
    rs->order = o;
    rs->cap = nc;
    return 0;
}

static uint64_t rs_push(rowset* rs, const orc_val* t, orc_ann a) {
    if (rs_reserve(rs, rs->n + 1)) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    uint64_t id = rs->n++;
    if (rs->arity) memcpy(rs->vals + id * rs->arity, t, rs->arity * sizeof(orc_val));
    rs->anns[id] = a;
    return id;
}

static inline const orc_val* rs_row(const rowset* rs, uint64_t id) {
    return rs->vals + id * rs->arity;
}

static int cmp_rowid(const void* pa, const void* pb, void* ctx) {
    const rowset* rs = (const rowset*)ctx;
    uint32_t a = *(const uint32_t*)pa, b = *(const uint32_t*)pb;
    return tup_cmp(rs_row(rs, a), rs_row(rs, b), rs->arity);
}

/* Bring the sorted view up to date: sort rows appended since the last call
 * and merge them into the existing order (rows are never removed). */
static void rs_sort(rowset* rs) {
    if (rs->order_n == rs->n) return;
    uint64_t old = rs->order_n, add = rs->n - old;
    for (uint64_t i = 0; i < add; i++) rs->order[old + i] = (uint32_t)(old + i);
    qsort_r(rs->order + old, add, sizeof(uint32_t), cmp_rowid, rs);
    if (old) {
        uint32_t* tmp = (uint32_t*)malloc(rs->n * sizeof(uint32_t));
        uint64_t i = 0, j = old, k = 0;
        while (i < old && j < rs->n) {
            if (tup_cmp(rs_row(rs, rs->order[j]), rs_row(rs, rs->order[i]), rs->arity) < 0)
                tmp[k++] = rs->order[j++];
            else
                tmp[k++] = rs->order[i++];
        }
        while (i < old) tmp[k++] = rs->order[i++];
        while (j < rs->n) tmp[k++] = rs->order[j++];
        memcpy(rs->order, tmp, rs->n * sizeof(uint32_t));
        free(tmp);
    }
    rs->order_n = rs->n;
}

typedef struct {
    const rowset* rs;
    uint32_t col;
} colctx;

static int cmp_colpos(const void* pa, const void* pb, void* ctx) {
    const colctx* c = (const colctx*)ctx;
    uint32_t a = *(const uint32_t*)pa, b = *(const uint32_t*)pb;
    const orc_val* va = rs_row(c->rs, c->rs->order[a]) + c->col;
    const orc_val* vb = rs_row(c->rs, c->rs->order[b]) + c->col;
    int r = val_cmp(va, vb);
    if (r) return r;
    return a < b ? -1 : (a > b);
}

static void rs_colidx(rowset* rs, uint32_t j) {
    rs_sort(rs);
    if (rs->colidx_n[j] == rs->n && rs->colidx[j]) return;
    free(rs->colidx[j]);
    rs->colidx[j] = (uint32_t*)malloc((rs->n ? rs->n : 1) * sizeof(uint32_t));
    for (uint64_t p = 0; p < rs->n; p++) rs->colidx[j][p] = (uint32_t)p;
    colctx c = {rs, j};
    qsort_r(rs->colidx[j], rs->n, sizeof(uint32_t), cmp_colpos, &c);
    rs->colidx_n[j] = rs->n;
}

/* Positions [lo, hi) of the sorted view whose column 0 equals v. */
static void rs_range0(const rowset* rs, const orc_val* v, uint64_t* lo, uint64_t* hi) {
    uint64_t a = 0, b = rs->n;
    while (a < b) {
        uint64_t m = (a + b) / 2;
        if (val_cmp(rs_row(rs, rs->order[m]), v) < 0) a = m + 1; else b = m;
    }
    *lo = a;
    b = rs->n;
    while (a < b) {
        uint64_t m = (a + b) / 2;
        if (val_cmp(rs_row(rs, rs->order[m]), v) <= 0) a = m + 1; else b = m;
    }
    *hi = a;
}

/* Index positions [lo, hi) of colidx[j] whose column j equals v. */
static void rs_rangej(const rowset* rs, uint32_t j, const orc_val* v, uint64_t* lo, uint64_t* hi) {
    const uint32_t* ix = rs->colidx[j];
    uint64_t a = 0, b = rs->n;
    while (a < b) {
        uint64_t m = (a + b) / 2;
        if (val_cmp(rs_row(rs, rs->order[ix[m]]) + j, v) < 0) a = m + 1; else b = m;
    }
    *lo = a;
    b = rs->n;
    while (a < b) {
        uint64_t m = (a + b) / 2;
        if (val_cmp(rs_row(rs, rs->order[ix[m]]) + j, v) <= 0) a = m + 1; else b = m;
    }
    *hi = a;
}

/* ---------------------------------------------------------------- relation */

typedef struct {
    char* name;
    uint32_t arity, semiring, stratum;
    rowset rs;
    uint64_t hcap;
    uint32_t* hslot;  /* row id + 1; 0 = empty */
    uint32_t* stamp;  /* round in which the row last changed */
    uint64_t stamp_cap;
} relation;

struct orc_db {
    const orc_program* P;
    uint32_t nrel;
    relation* rel;
    int32_t* atom_rel;     /* atom index -> relation index, -1 if undeclared */
    int32_t* atom_strat;   /* stratum of the atom's predicate (pred_stratum) or -1 */
    uint64_t* iters;
    uint64_t universe;
    uint64_t steps;
    uint32_t round;
    int32_t pending;       /* an EDB-load error to report from orc_run */
    orc_error pending_err;
};

static void rel_rehash(relation* r, uint64_t ncap) {
    uint32_t* hs = (uint32_t*)calloc(ncap, sizeof(uint32_t));
    for (uint64_t id = 0; id < r->rs.n; id++) {
        uint64_t h = tup_hash(rs_row(&r->rs, id), r->arity) & (ncap - 1);
        while (hs[h]) h = (h + 1) & (ncap - 1);
        hs[h] = (uint32_t)(id + 1);
    }
    free(r->hslot);
    r->hslot = hs;
    r->hcap = ncap;
}

static int64_t rel_find(const relation* r, const orc_val* t) {
    if (!r->hcap) return -1;
    uint64_t h = tup_hash(t, r->arity) & (r->hcap - 1);
    while (r->hslot[h]) {
        uint64_t id = r->hslot[h] - 1;
        if (tup_cmp(rs_row(&r->rs, id), t, r->arity) == 0) return (int64_t)id;
        h = (h + 1) & (r->hcap - 1);
    }
    return -1;
}

/* Relation::combine (store.rs:34-48). */
static int rel_combine(orc_db* db, relation* r, const orc_val* t, orc_ann a, int* changed) {
    int64_t id = rel_find(r, t);
    if (id < 0) {
        if ((r->rs.n + 1) * 2 > r->hcap) rel_rehash(r, r->hcap ? r->hcap * 2 : 1024);
        uint64_t nid = rs_push(&r->rs, t, a);
        if (r->stamp_cap < r->rs.cap) {
            r->stamp = (uint32_t*)realloc(r->stamp, r->rs.cap * sizeof(uint32_t));
            r->stamp_cap = r->rs.cap;
        }
        r->stamp[nid] = db->round;
        uint64_t h = tup_hash(t, r->arity) & (r->hcap - 1);
        while (r->hslot[h]) h = (h + 1) & (r->hcap - 1);
        r->hslot[h] = (uint32_t)(nid + 1);
        *changed = 1;
        return ORC_OK;
    }
    orc_ann merged;
    int ch = 0;
    if (ann_oplus(r->rs.anns[id], a, &merged, &ch) != ORC_OK) return ORC_PANIC;
    if (ch) {
        r->rs.anns[id] = merged;
        r->stamp[id] = db->round;
    }
    *changed = ch;
    return ORC_OK;
}

static int32_t find_rel(const orc_db* db, const char* name) {
    /* Db is a BTreeMap by name; a repeated declaration keeps the last one. */
    for (int32_t i = (int32_t)db->nrel - 1; i >= 0; i--)
        if (strcmp(db->rel[i].name, name) == 0) return i;
    return -1;
}

static void set_err(orc_error* err, int code, const char* pred, const char* msg) {
    if (!err) return;
    err->code = code;
    if (pred) snprintf(err->pred, sizeof err->pred, "%s", pred);
    if (msg) snprintf(err->msg, sizeof err->msg, "%s", msg);
}

orc_db* orc_db_new(const orc_program* P) {
    orc_db* db = (orc_db*)calloc(1, sizeof *db);
    db->P = P;
    db->nrel = P->n_preds;
    db->rel = (relation*)calloc(P->n_preds ? P->n_preds : 1, sizeof(relation));
    for (uint32_t i = 0; i < P->n_preds; i++) {
        relation* r = &db->rel[i];
        r->name = strdup(P->preds[i].name);
        r->arity = P->preds[i].arity;
        r->semiring = P->preds[i].semiring;
        r->stratum = P->preds[i].stratum;
        rs_init(&r->rs, r->arity);
    }
    db->atom_rel = (int32_t*)malloc((P->n_atoms ? P->n_atoms : 1) * sizeof(int32_t));
    db->atom_strat = (int32_t*)malloc((P->n_atoms ? P->n_atoms : 1) * sizeof(int32_t));
    for (uint32_t a = 0; a < P->n_atoms; a++) {
        int32_t ri = find_rel(db, P->atoms[a].pred);
        db->atom_rel[a] = ri;
        db->atom_strat[a] = ri >= 0 ? (int32_t)db->rel[ri].stratum : -1;
    }
    db->iters = (uint64_t*)calloc(P->num_strata ? P->num_strata : 1, sizeof(uint64_t));
    return db;
}

void orc_db_free(orc_db* db) {
    if (!db) return;
    for (uint32_t i = 0; i < db->nrel; i++) {
        free(db->rel[i].name);
        rs_free(&db->rel[i].rs);
        free(db->rel[i].hslot);
        free(db->rel[i].stamp);
    }
    free(db->rel);
    free(db->atom_rel);
    free(db->atom_strat);
    free(db->iters);
    free(db);
}

static int32_t insert_common(orc_db* db, const char* pred, int32_t* ri, orc_error* err) {
    if (db->pending) return db->pending;
    *ri = find_rel(db, pred);
    if (*ri < 0) {
        db->pending = ORC_ERR_UNKNOWN_PRED;
        set_err(&db->pending_err, ORC_ERR_UNKNOWN_PRED, pred, "unknown predicate");
        if (err) *err = db->pending_err;
        return db->pending;
    }
    return ORC_OK;
}

int32_t orc_insert(orc_db* db, const char* pred, uint64_t n, const orc_val* rows,
                   const orc_ann* anns, orc_error* err) {
    int32_t ri;
    int32_t s = insert_common(db, pred, &ri, err);
    if (s) return s;
    relation* r = &db->rel[ri];
    for (uint64_t i = 0; i < n; i++) {
        int ch;
        orc_ann a = anns ? anns[i] : ann_unit();
        if (rel_combine(db, r, rows + i * r->arity, a, &ch) != ORC_OK) {
            db->pending = ORC_PANIC;
            set_err(&db->pending_err, ORC_PANIC, pred, "semiring annotation mismatch in oplus");
            if (err) *err = db->pending_err;
            return ORC_PANIC;
        }
    }
    return ORC_OK;
}

int32_t orc_insert_sym(orc_db* db, const char* pred, uint64_t n, uint32_t arity,
                       const uint32_t* const* cols, const int64_t* w,
                       const uint8_t* inf, orc_error* err) {
    int32_t ri;
    int32_t s = insert_common(db, pred, &ri, err);
    if (s) return s;
    relation* r = &db->rel[ri];
    if (arity != r->arity) {
        db->pending = ORC_ERR_BAD_INPUT;
        set_err(&db->pending_err, ORC_ERR_BAD_INPUT, pred, "EDB arity differs from the declaration");
        if (err) *err = db->pending_err;
        return ORC_ERR_BAD_INPUT;
    }
    orc_val t[16];
    if (arity > 16) return ORC_ERR_BAD_INPUT;
    for (uint64_t i = 0; i < n; i++) {
        for (uint32_t j = 0; j < arity; j++) {
            t[j].kind = ORC_SYM;
            t[j].pad = 0;
            t[j].v = cols[j][i];
        }
        orc_ann a = ann_unit();
        if (w) {
            a = ann_w(w[i]);
            if (inf && inf[i]) {
                a.inf = 1;
                a.w = 0;
            }
        }
        int ch;
        if (rel_combine(db, r, t, a, &ch) != ORC_OK) {
            db->pending = ORC_PANIC;
            set_err(&db->pending_err, ORC_PANIC, pred, "semiring annotation mismatch in oplus");
            if (err) *err = db->pending_err;
            return ORC_PANIC;
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------- solve */

typedef struct sctx sctx;
typedef int (*emit_fn)(sctx* c, const orc_val* binding, orc_ann acc);

struct sctx {
    orc_db* db;
    const orc_program* P;
    uint32_t cur;           /* stratum being saturated */
    uint32_t body_first;    /* first body atom index */
    uint32_t nbody;
    int dpos;               /* delta literal position, -1 for none */
    rowset* drows;          /* its rows */
    int indexed;
    uint32_t nvars;
    emit_fn emit;
    void* ud;
    orc_error* err;
};

/* unify (naive.rs:407-455) of one atom against one row.  `bound` marks the
 * bound slots; on failure the caller restores the binding. */
static int unify(const orc_program* P, const orc_atom* at, const orc_val* row, uint32_t arity,
                 orc_val* b, uint8_t* bound) {
    if (at->n_terms != arity) return 0;
    for (uint32_t k = 0; k < at->n_terms; k++) {
        const orc_term* t = &P->terms[at->first_term + k];
        const orc_val* v = row + k;
        switch (t->kind) {
        case ORC_T_VAR:
            if (bound[t->a]) {
                if (!val_eq(&b[t->a], v)) return 0;
            } else {
                b[t->a] = *v;
                bound[t->a] = 1;
            }
            break;
        case ORC_T_CONST:
            if (!(v->kind == ORC_SYM && v->v == (int64_t)t->a)) return 0;
            break;
        case ORC_T_INT:
            if (!(v->kind == ORC_INT && v->v == t->b)) return 0;
            break;
        default: /* Agg never unifies; a compound needs a Term value */
            return 0;
        }
    }
    return 1;
}

/* The first argument whose value is already determined before the scan,
 * used to narrow the scan without changing which rows unify or their order. */
static int determined_col(const orc_program* P, const orc_atom* at, const orc_val* b,
                          const uint8_t* bound, orc_val* out) {
    for (uint32_t k = 0; k < at->n_terms; k++) {
        const orc_term* t = &P->terms[at->first_term + k];
        if (t->kind == ORC_T_VAR && bound[t->a]) {
            *out = b[t->a];
            return (int)k;
        }
        if (t->kind == ORC_T_CONST) {
            out->kind = ORC_SYM; out->pad = 0; out->v = t->a;
            return (int)k;
        }
        if (t->kind == ORC_T_INT) {
            out->kind = ORC_INT; out->pad = 0; out->v = t->b;
            return (int)k;
        }
    }
    return -1;
}

/* ground_term / ground_atom (naive.rs:461-511) for negated atoms. */
static int ground_neg(sctx* c, const orc_atom* at, const orc_val* b, const uint8_t* bound,
                      orc_val* out) {
    for (uint32_t k = 0; k < at->n_terms; k++) {
        const orc_term* t = &c->P->terms[at->first_term + k];
        switch (t->kind) {
        case ORC_T_VAR:
            if (!bound[t->a]) {
                c->err->code = ORC_ERR_UNSAFE_NEG_VAR;
                c->err->slot = t->a;
                snprintf(c->err->msg, sizeof c->err->msg,
                         "unbound variable in negated literal, slot %u", t->a);
                return ORC_ERR_UNSAFE_NEG_VAR;
            }
            out[k] = b[t->a];
            break;
        case ORC_T_CONST:
            out[k].kind = ORC_SYM; out[k].pad = 0; out[k].v = t->a;
            break;
        case ORC_T_INT:
            out[k].kind = ORC_INT; out[k].pad = 0; out[k].v = t->b;
            break;
        case ORC_T_AGG:
            set_err(c->err, ORC_ERR_UNSUPPORTED, NULL, "aggregate in atom position");
            return ORC_ERR_UNSUPPORTED;
        default:
            set_err(c->err, ORC_ERR_UNSUPPORTED, NULL, "compound terms are not evaluated by this oracle");
            return ORC_ERR_UNSUPPORTED;
        }
    }
    return ORC_OK;
}

static int solve(sctx* c, uint32_t idx, orc_val* b, uint8_t* bound, orc_ann acc);

static int scan_rows(sctx* c, uint32_t idx, const orc_atom* at, rowset* rs, orc_val* b,
                     uint8_t* bound, orc_ann acc) {
    uint32_t nv = c->nvars;
    orc_val* nb = (orc_val*)alloca((nv ? nv : 1) * sizeof(orc_val));
    uint8_t* nbd = (uint8_t*)alloca(nv ? nv : 1);
    rs_sort(rs);
    uint64_t lo = 0, hi = rs->n;
    int col = -1;
    orc_val key;
    if (c->indexed && at->n_terms == rs->arity) {
        col = determined_col(c->P, at, b, bound, &key);
        if (col == 0) {
            rs_range0(rs, &key, &lo, &hi);
        } else if (col > 0) {
            rs_colidx(rs, (uint32_t)col);
            rs_rangej(rs, (uint32_t)col, &key, &lo, &hi);
        }
    }
    for (uint64_t p = lo; p < hi; p++) {
        uint64_t id = col > 0 ? rs->order[rs->colidx[col][p]] : rs->order[p];
        memcpy(nb, b, nv * sizeof(orc_val));
        memcpy(nbd, bound, nv);
        c->db->steps++;
        if (unify(c->P, at, rs_row(rs, id), rs->arity, nb, nbd)) {
            orc_ann na;
            int s = ann_otimes(acc, rs->anns[id], &na, c->err);
            if (s) return s;
            s = solve(c, idx + 1, nb, nbd, na);
            if (s) return s;
        }
    }
    return ORC_OK;
}

/* solve (seminaive.rs:215-298; naive.rs:237-297). */
static int solve(sctx* c, uint32_t idx, orc_val* b, uint8_t* bound, orc_ann acc) {
    if (idx == c->nbody) return c->emit(c, b, acc);
    uint32_t ai = c->body_first + idx;
    const orc_atom* at = &c->P->atoms[ai];
    if (!at->negated) {
        if (c->dpos == (int)idx) return scan_rows(c, idx, at, c->drows, b, bound, acc);
        int32_t ri = c->db->atom_rel[ai];
        if (ri < 0) {
            set_err(c->err, ORC_ERR_UNKNOWN_PRED, at->pred, "unknown predicate");
            return ORC_ERR_UNKNOWN_PRED;
        }
        return scan_rows(c, idx, at, &c->db->rel[ri].rs, b, bound, acc);
    }
    /* negation_holds (naive.rs:537-561) */
    int32_t ns = c->db->atom_strat[ai];
    if (ns < 0) {
        set_err(c->err, ORC_ERR_UNKNOWN_PRED, at->pred, "unknown predicate");
        return ORC_ERR_UNKNOWN_PRED;
    }
    if ((uint32_t)ns >= c->cur) {
        set_err(c->err, ORC_ERR_NEGATION_NOT_STRATIFIED, at->pred,
                "negated predicate is not in a strictly lower stratum");
        return ORC_ERR_NEGATION_NOT_STRATIFIED;
    }
    orc_val t[16];
    if (at->n_terms > 16) {
        set_err(c->err, ORC_ERR_BAD_INPUT, at->pred, "atom arity above 16");
        return ORC_ERR_BAD_INPUT;
    }
    int s = ground_neg(c, at, b, bound, t);
    if (s) return s;
    relation* r = &c->db->rel[c->db->atom_rel[ai]];
    int present = (at->n_terms == r->arity) && rel_find(r, t) >= 0;
    if (present) return ORC_OK;
    return solve(c, idx + 1, b, bound, acc);
}

/* -------------------------------------------------- head building / rounds */

typedef struct {
    uint32_t nrel;
    rowset* derived;   /* per relation, this round's derivations in order */
    const orc_atom* head;
    int32_t head_rel;
    int head_err;      /* first failing head argument's error, 0 if none */
    uint32_t head_err_slot;
    uint64_t nsol;
} round_acc;

static int emit_head(sctx* c, const orc_val* b, orc_ann acc) {
    round_acc* ra = (round_acc*)c->ud;
    ra->nsol++;
    if (ra->head_err) return ORC_OK;  /* raised once the rule's solve is done */
    const orc_atom* h = ra->head;
    orc_val t[16];
    for (uint32_t k = 0; k < h->n_terms; k++) {
        const orc_term* tm = &c->P->terms[h->first_term + k];
        if (tm->kind == ORC_T_VAR) t[k] = b[tm->a];
        else if (tm->kind == ORC_T_CONST) { t[k].kind = ORC_SYM; t[k].pad = 0; t[k].v = tm->a; }
        else { t[k].kind = ORC_INT; t[k].pad = 0; t[k].v = tm->b; }
    }
    rs_push(&ra->derived[ra->head_rel], t, acc);
    return ORC_OK;
}

/* Static part of build_head (naive.rs:526-532): a head variable is bound in
 * every solution iff it occurs in a positive body literal. */
static void head_check(const orc_program* P, const orc_rule* r, int* err, uint32_t* slot) {
    *err = 0;
    const orc_atom* h = &P->atoms[r->head];
    for (uint32_t k = 0; k < h->n_terms; k++) {
        const orc_term* tm = &P->terms[h->first_term + k];
        if (tm->kind == ORC_T_VAR) {
            int seen = 0;
            for (uint32_t l = 0; l < r->n_lits && !seen; l++) {
                const orc_atom* at = &P->atoms[r->first_lit + l];
                if (at->negated) continue;
                for (uint32_t q = 0; q < at->n_terms; q++) {
                    const orc_term* bt = &P->terms[at->first_term + q];
                    if (bt->kind == ORC_T_VAR && bt->a == tm->a) { seen = 1; break; }
                }
            }
            if (!seen) { *err = ORC_ERR_UNSAFE_HEAD_VAR; *slot = tm->a; return; }
        } else if (tm->kind == ORC_T_AGG || tm->kind == ORC_T_COMPOUND) {
            *err = ORC_ERR_UNSUPPORTED;
            return;
        }
    }
}

static int is_agg_rule(const orc_program* P, const orc_rule* r) {
    const orc_atom* h = &P->atoms[r->head];
    for (uint32_t k = 0; k < h->n_terms; k++)
        if (P->terms[h->first_term + k].kind == ORC_T_AGG) return 1;
    return 0;
}

/* Solve one rule (optionally with a delta literal) and collect its heads. */
static int solve_rule(orc_db* db, const orc_rule* r, int dpos, rowset* drows, int indexed,
                      round_acc* ra, orc_error* err) {
    const orc_program* P = db->P;
    int32_t hr = find_rel(db, P->atoms[r->head].pred);
    if (hr < 0) {
        set_err(err, ORC_ERR_UNKNOWN_PRED, P->atoms[r->head].pred, "unknown predicate");
        return ORC_ERR_UNKNOWN_PRED;
    }
    if (P->atoms[r->head].n_terms != db->rel[hr].arity) {
        set_err(err, ORC_ERR_BAD_INPUT, P->atoms[r->head].pred,
                "head arity differs from the declaration");
        return ORC_ERR_BAD_INPUT;
    }
    ra->head = &P->atoms[r->head];
    ra->head_rel = hr;
    ra->nsol = 0;
    uint32_t hslot = 0;
    head_check(P, r, &ra->head_err, &hslot);
    ra->head_err_slot = hslot;
    sctx c;
    memset(&c, 0, sizeof c);
    c.db = db; c.P = P; c.cur = r->stratum; c.body_first = r->first_lit; c.nbody = r->n_lits;
    c.dpos = dpos; c.drows = drows; c.indexed = indexed; c.nvars = r->var_count;
    c.emit = emit_head; c.ud = ra; c.err = err;
    orc_val* b = (orc_val*)calloc(r->var_count ? r->var_count : 1, sizeof(orc_val));
    uint8_t* bound = (uint8_t*)calloc(r->var_count ? r->var_count : 1, 1);
    int s = solve(&c, 0, b, bound, ann_otimes_id(db->rel[hr].semiring));
    free(b);
    free(bound);
    if (s) return s;
    if (ra->nsol && ra->head_err) {
        if (ra->head_err == ORC_ERR_UNSAFE_HEAD_VAR) {
            err->code = ORC_ERR_UNSAFE_HEAD_VAR;
            err->slot = ra->head_err_slot;
            snprintf(err->msg, sizeof err->msg, "unbound head variable in slot %u", ra->head_err_slot);
        } else {
            set_err(err, ORC_ERR_UNSUPPORTED, NULL, "unsupported head term");
        }
        return ra->head_err;
    }
    return ORC_OK;
}

/* Apply the round's derivations (Jacobi), stamping changed rows. */
static int apply_derived(orc_db* db, round_acc* ra, int* any_changed, orc_error* err) {
    *any_changed = 0;
    for (uint32_t i = 0; i < ra->nrel; i++) {
        rowset* d = &ra->derived[i];
        relation* r = &db->rel[i];
        for (uint64_t k = 0; k < d->n; k++) {
            int ch;
            if (rel_combine(db, r, rs_row(d, k), d->anns[k], &ch) != ORC_OK) {
                set_err(err, ORC_PANIC, r->name, "semiring annotation mismatch in oplus");
                return ORC_PANIC;
            }
            *any_changed |= ch;
        }
    }
    return ORC_OK;
}

static void ra_init(round_acc* ra, orc_db* db) {
    memset(ra, 0, sizeof *ra);
    ra->nrel = db->nrel;
    ra->derived = (rowset*)calloc(db->nrel ? db->nrel : 1, sizeof(rowset));
    for (uint32_t i = 0; i < db->nrel; i++) rs_init(&ra->derived[i], db->rel[i].arity);
}

static void ra_free(round_acc* ra) {
    for (uint32_t i = 0; i < ra->nrel; i++) rs_free(&ra->derived[i]);
    free(ra->derived);
}

/* -------------------------------------------------------------- aggregates */

typedef struct {
    orc_val* keys;   /* nsol x klen */
    orc_val* vals;
    uint64_t n, cap;
    uint32_t klen;
    uint32_t agg_index, agg_slot;
    const orc_atom* head;
    int err;
    uint32_t err_slot;
} agg_acc;

static int emit_agg(sctx* c, const orc_val* b, orc_ann acc) {
    (void)acc;
    agg_acc* g = (agg_acc*)c->ud;
    if (g->n == g->cap) {
        g->cap = g->cap ? g->cap * 2 : 256;
        g->keys = (orc_val*)realloc(g->keys, g->cap * (g->klen ? g->klen : 1) * sizeof(orc_val));
        g->vals = (orc_val*)realloc(g->vals, g->cap * sizeof(orc_val));
    }
    g->n++;
    uint32_t kk = 0;
    for (uint32_t k = 0; k < g->head->n_terms; k++) {
        if (k == g->agg_index) continue;
        const orc_term* tm = &c->P->terms[g->head->first_term + k];
        orc_val v;
        if (tm->kind == ORC_T_VAR) v = b[tm->a];
        else if (tm->kind == ORC_T_CONST) { v.kind = ORC_SYM; v.pad = 0; v.v = tm->a; }
        else { v.kind = ORC_INT; v.pad = 0; v.v = tm->b; }
        g->keys[(g->n - 1) * g->klen + kk++] = v;
    }
    g->vals[g->n - 1] = b[g->agg_slot];
    return ORC_OK;
}

typedef struct {
    const agg_acc* g;
} aggsort_ctx;

static int cmp_aggrow(const void* pa, const void* pb, void* ctx) {
    const agg_acc* g = (const agg_acc*)ctx;
    uint64_t a = *(const uint64_t*)pa, b = *(const uint64_t*)pb;
    int r = tup_cmp(g->keys + a * g->klen, g->keys + b * g->klen, g->klen);
    if (r) return r;
    return a < b ? -1 : (a > b);
}

/* eval_aggregate_rule + fold_aggregate (naive.rs:300-403). */
static int eval_aggregate(orc_db* db, const orc_rule* r, int indexed, orc_error* err) {
    const orc_program* P = db->P;
    for (uint32_t l = 0; l < r->n_lits; l++) {
        uint32_t ai = r->first_lit + l;
        int32_t s = db->atom_strat[ai];
        if (s < 0) {
            set_err(err, ORC_ERR_UNKNOWN_PRED, P->atoms[ai].pred, "unknown predicate");
            return ORC_ERR_UNKNOWN_PRED;
        }
        if ((uint32_t)s >= r->stratum) {
            set_err(err, ORC_ERR_AGGREGATE_IN_RECURSION, P->atoms[ai].pred,
                    "aggregate body predicate is not in a strictly lower stratum");
            return ORC_ERR_AGGREGATE_IN_RECURSION;
        }
    }
    const orc_atom* h = &P->atoms[r->head];
    int found = 0;
    uint32_t agg_index = 0, op = 0, slot = 0;
    for (uint32_t k = 0; k < h->n_terms; k++) {
        const orc_term* tm = &P->terms[h->first_term + k];
        if (tm->kind == ORC_T_AGG) {
            if (found) {
                set_err(err, ORC_ERR_MALFORMED_AGGREGATE_HEAD, NULL,
                        "aggregate head must contain exactly one agg term");
                return ORC_ERR_MALFORMED_AGGREGATE_HEAD;
            }
            found = 1;
            agg_index = k;
            op = tm->a;
            slot = (uint32_t)tm->b;
        }
    }
    if (!found) {
        set_err(err, ORC_ERR_MALFORMED_AGGREGATE_HEAD, NULL,
                "aggregate head must contain exactly one agg term");
        return ORC_ERR_MALFORMED_AGGREGATE_HEAD;
    }
    if (slot >= r->var_count) {
        set_err(err, ORC_PANIC, NULL, "aggregate slot out of range");
        return ORC_PANIC;
    }
    agg_acc g;
    memset(&g, 0, sizeof g);
    g.klen = h->n_terms - 1;
    g.agg_index = agg_index;
    g.agg_slot = slot;
    g.head = h;
    sctx c;
    memset(&c, 0, sizeof c);
    c.db = db; c.P = P; c.cur = r->stratum; c.body_first = r->first_lit; c.nbody = r->n_lits;
    c.dpos = -1; c.indexed = indexed; c.nvars = r->var_count;
    c.emit = emit_agg; c.ud = &g; c.err = err;
    orc_val* b = (orc_val*)calloc(r->var_count ? r->var_count : 1, sizeof(orc_val));
    uint8_t* bound = (uint8_t*)calloc(r->var_count ? r->var_count : 1, 1);
    int s = solve(&c, 0, b, bound, ann_unit());
    free(b);
    free(bound);
    if (s) { free(g.keys); free(g.vals); return s; }
    if (g.n) {
        /* head_arg_value and the aggregand must be bound (naive.rs:350-362) */
        int herr = 0;
        uint32_t hslot = 0;
        for (uint32_t k = 0; k < h->n_terms && !herr; k++) {
            if (k == agg_index) continue;
            const orc_term* tm = &P->terms[h->first_term + k];
            if (tm->kind == ORC_T_VAR) {
                int seen = 0;
                for (uint32_t l = 0; l < r->n_lits && !seen; l++) {
                    const orc_atom* at = &P->atoms[r->first_lit + l];
                    if (at->negated) continue;
                    for (uint32_t q = 0; q < at->n_terms; q++) {
                        const orc_term* bt = &P->terms[at->first_term + q];
                        if (bt->kind == ORC_T_VAR && bt->a == tm->a) { seen = 1; break; }
                    }
                }
                if (!seen) { herr = ORC_ERR_UNSAFE_HEAD_VAR; hslot = tm->a; }
            } else if (tm->kind == ORC_T_COMPOUND) {
                herr = ORC_ERR_UNSUPPORTED;
            }
        }
        if (!herr) {
            int seen = 0;
            for (uint32_t l = 0; l < r->n_lits && !seen; l++) {
                const orc_atom* at = &P->atoms[r->first_lit + l];
                if (at->negated) continue;
                for (uint32_t q = 0; q < at->n_terms; q++) {
                    const orc_term* bt = &P->terms[at->first_term + q];
                    if (bt->kind == ORC_T_VAR && bt->a == slot) { seen = 1; break; }
                }
            }
            if (!seen) { herr = ORC_ERR_UNSAFE_HEAD_VAR; hslot = slot; }
        }
        if (herr) {
            if (herr == ORC_ERR_UNSAFE_HEAD_VAR) {
                err->code = herr;
                err->slot = hslot;
                snprintf(err->msg, sizeof err->msg, "unbound head variable in slot %u", hslot);
            } else {
                set_err(err, herr, NULL, "unsupported head term");
            }
            free(g.keys); free(g.vals);
            return herr;
        }
    }
    /* group by key (HashMap in the reference; group order does not affect
     * the result), fold each group */
    uint64_t* ord = (uint64_t*)malloc((g.n ? g.n : 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < g.n; i++) ord[i] = i;
    qsort_r(ord, g.n, sizeof(uint64_t), cmp_aggrow, &g);
    int32_t hr = find_rel(db, h->pred);
    int rc = ORC_OK;
    for (uint64_t i = 0; i < g.n && rc == ORC_OK;) {
        uint64_t j = i + 1;
        while (j < g.n && tup_cmp(g.keys + ord[i] * g.klen, g.keys + ord[j] * g.klen, g.klen) == 0) j++;
        int64_t res = 0;
        if (op == ORC_AGG_COUNT) {
            res = (int64_t)(j - i);
        } else if (op == ORC_AGG_PROB_OR) {
            set_err(err, ORC_ERR_UNSUPPORTED, NULL, "prob-or aggregate");
            rc = ORC_ERR_UNSUPPORTED;
            break;
        } else {
            for (uint64_t k = i; k < j; k++) {
                const orc_val* v = &g.vals[ord[k]];
                if (v->kind != ORC_INT) {
                    set_err(err, ORC_ERR_NON_INTEGER_AGGREGAND, NULL,
                            "min/max/sum require an integer aggregand");
                    rc = ORC_ERR_NON_INTEGER_AGGREGAND;
                    break;
                }
                if (k == i) res = op == ORC_AGG_SUM ? 0 : v->v;
                if (op == ORC_AGG_SUM) res = (int64_t)((uint64_t)res + (uint64_t)v->v);
                else if (op == ORC_AGG_MIN) { if (v->v < res) res = v->v; }
                else if (op == ORC_AGG_MAX) { if (v->v > res) res = v->v; }
            }
            if (rc) break;
        }
        if (hr < 0) {
            set_err(err, ORC_PANIC, h->pred, "insert into unknown relation");
            rc = ORC_PANIC;
            break;
        }
        orc_val t[16];
        uint32_t kk = 0;
        for (uint32_t k = 0; k < h->n_terms; k++) {
            if (k == agg_index) { t[k].kind = ORC_INT; t[k].pad = 0; t[k].v = res; }
            else t[k] = g.keys[ord[i] * g.klen + kk++];
        }
        int ch;
        if (rel_combine(db, &db->rel[hr], t, ann_unit(), &ch) != ORC_OK) {
            set_err(err, ORC_PANIC, h->pred, "semiring annotation mismatch in oplus");
            rc = ORC_PANIC;
        }
        i = j;
    }
    free(ord);
    free(g.keys);
    free(g.vals);
    return rc;
}

/* --------------------------------------------------------------- strata */

static int stratum_is_trop(const orc_db* db, uint32_t k) {
    const orc_program* P = db->P;
    for (uint32_t i = 0; i < P->n_rules; i++) {
        const orc_rule* r = &P->rules[i];
        if (r->stratum != k) continue;
        int32_t hr = find_rel(db, P->atoms[r->head].pred);
        if (hr >= 0 && db->rel[hr].semiring == ORC_TROP) return 1;
    }
    return 0;
}

static int cmp64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : (x > y);
}

static uint64_t distinct_symbols(orc_db* db) {
    uint64_t total = 0;
    for (uint32_t i = 0; i < db->nrel; i++) total += db->rel[i].rs.n * db->rel[i].arity;
    int64_t* syms = (int64_t*)malloc((total ? total : 1) * sizeof(int64_t));
    uint64_t m = 0;
    for (uint32_t i = 0; i < db->nrel; i++) {
        const rowset* rs = &db->rel[i].rs;
        for (uint64_t k = 0; k < rs->n * rs->arity; k++)
            if (rs->vals[k].kind == ORC_SYM) syms[m++] = rs->vals[k].v;
    }
    qsort(syms, m, sizeof(int64_t), cmp64);
    uint64_t u = 0;
    for (uint64_t k = 0; k < m; k++)
        if (k == 0 || syms[k] != syms[k - 1]) u++;
    free(syms);
    return u;
}

/* Build per-relation delta rowsets from rows stamped in the current round. */
static void build_delta(orc_db* db, rowset* delta, const uint8_t* is_idb) {
    for (uint32_t i = 0; i < db->nrel; i++) {
        rs_free(&delta[i]);
        rs_init(&delta[i], db->rel[i].arity);
        if (!is_idb[i]) continue;
        relation* r = &db->rel[i];
        for (uint64_t id = 0; id < r->rs.n; id++)
            if (r->stamp[id] == db->round) rs_push(&delta[i], rs_row(&r->rs, id), r->rs.anns[id]);
    }
}

static int saturate_semi(orc_db* db, uint32_t k, int indexed, orc_error* err) {
    const orc_program* P = db->P;
    for (uint32_t i = 0; i < P->n_rules; i++) {
        const orc_rule* r = &P->rules[i];
        if (r->stratum == k && is_agg_rule(P, r)) {
            int s = eval_aggregate(db, r, indexed, err);
            if (s) return s;
        }
    }
    uint8_t* is_idb = (uint8_t*)calloc(db->nrel ? db->nrel : 1, 1);
    for (uint32_t i = 0; i < P->n_rules; i++) {
        const orc_rule* r = &P->rules[i];
        if (r->stratum != k || is_agg_rule(P, r)) continue;
        int32_t hr = find_rel(db, P->atoms[r->head].pred);
        if (hr >= 0) is_idb[hr] = 1;
    }
    rowset* delta = (rowset*)calloc(db->nrel ? db->nrel : 1, sizeof(rowset));
    for (uint32_t i = 0; i < db->nrel; i++) rs_init(&delta[i], db->rel[i].arity);
    int rc = ORC_OK;
    int is_trop = stratum_is_trop(db, k);
    uint64_t cap = db->universe + 2, iters = 0;
    int bootstrap = 1;
    for (;;) {
        round_acc ra;
        ra_init(&ra, db);
        db->round++;
        for (uint32_t i = 0; i < P->n_rules && rc == ORC_OK; i++) {
            const orc_rule* r = &P->rules[i];
            if (r->stratum != k || is_agg_rule(P, r)) continue;
            if (bootstrap) {
                rc = solve_rule(db, r, -1, NULL, indexed, &ra, err);
                continue;
            }
            /* a head lookup precedes the delta positions (seminaive.rs:128-131) */
            if (find_rel(db, P->atoms[r->head].pred) < 0) {
                set_err(err, ORC_ERR_UNKNOWN_PRED, P->atoms[r->head].pred, "unknown predicate");
                rc = ORC_ERR_UNKNOWN_PRED;
                break;
            }
            for (uint32_t l = 0; l < r->n_lits && rc == ORC_OK; l++) {
                uint32_t ai = r->first_lit + l;
                if (P->atoms[ai].negated) continue;
                int32_t ri = db->atom_rel[ai];
                if (ri < 0 || !is_idb[ri]) continue;
                if (delta[ri].n == 0) continue;
                rc = solve_rule(db, r, (int)l, &delta[ri], indexed, &ra, err);
            }
        }
        if (rc == ORC_OK) {
            int any;
            rc = apply_derived(db, &ra, &any, err);
        }
        ra_free(&ra);
        if (rc) break;
        build_delta(db, delta, is_idb);
        if (!bootstrap) {
            iters++;
            if (is_trop && iters > cap) {
                set_err(err, ORC_ERR_NEGATIVE_WEIGHT_CYCLE, NULL,
                        "negative-weight cycle: tropical stratum diverges");
                rc = ORC_ERR_NEGATIVE_WEIGHT_CYCLE;
                break;
            }
        }
        bootstrap = 0;
        int nonempty = 0;
        for (uint32_t i = 0; i < db->nrel; i++)
            if (delta[i].n) nonempty = 1;
        if (!nonempty) break;
    }
    db->iters[k] = iters;
    for (uint32_t i = 0; i < db->nrel; i++) rs_free(&delta[i]);
    free(delta);
    free(is_idb);
    return rc;
}

static int saturate_naive(orc_db* db, uint32_t k, int indexed, orc_error* err) {
    const orc_program* P = db->P;
    for (uint32_t i = 0; i < P->n_rules; i++) {
        const orc_rule* r = &P->rules[i];
        if (r->stratum == k && is_agg_rule(P, r)) {
            int s = eval_aggregate(db, r, indexed, err);
            if (s) return s;
        }
    }
    int is_trop = stratum_is_trop(db, k);
    uint64_t cap = db->universe + 2, iters = 0;
    int rc = ORC_OK;
    for (;;) {
        round_acc ra;
        ra_init(&ra, db);
        db->round++;
        for (uint32_t i = 0; i < P->n_rules && rc == ORC_OK; i++) {
            const orc_rule* r = &P->rules[i];
            if (r->stratum != k || is_agg_rule(P, r)) continue;
            rc = solve_rule(db, r, -1, NULL, indexed, &ra, err);
        }
        int changed = 0;
        if (rc == ORC_OK) rc = apply_derived(db, &ra, &changed, err);
        ra_free(&ra);
        if (rc) break;
        iters++;
        if (!changed) break;
        if (is_trop && iters > cap) {
            set_err(err, ORC_ERR_NEGATIVE_WEIGHT_CYCLE, NULL,
                    "negative-weight cycle: tropical stratum diverges");
            rc = ORC_ERR_NEGATIVE_WEIGHT_CYCLE;
            break;
        }
    }
    db->iters[k] = iters;
    return rc;
}

int32_t orc_run(orc_db* db, int semi, int indexed, orc_error* err) {
    orc_error local;
    if (!err) err = &local;
    memset(err, 0, sizeof *err);
    if (db->pending) {
        *err = db->pending_err;
        return db->pending;
    }
    db->universe = distinct_symbols(db);
    for (uint32_t k = 0; k < db->P->num_strata; k++) {
        int s = semi ? saturate_semi(db, k, indexed, err) : saturate_naive(db, k, indexed, err);
        if (s) {
            err->code = s;
            return s;
        }
    }
    return ORC_OK;
}

uint64_t orc_iterations(const orc_db* db, uint32_t stratum) {
    return stratum < db->P->num_strata ? db->iters[stratum] : 0;
}

uint64_t orc_universe(const orc_db* db) { return db->universe; }

uint64_t orc_unify_steps(const orc_db* db) { return db->steps; }

int64_t orc_rel_size(orc_db* db, const char* pred) {
    int32_t ri = find_rel(db, pred);
    return ri < 0 ? -1 : (int64_t)db->rel[ri].rs.n;
}

int64_t orc_rel_arity(orc_db* db, const char* pred) {
    int32_t ri = find_rel(db, pred);
    return ri < 0 ? -1 : (int64_t)db->rel[ri].arity;
}

int32_t orc_rel_copy(orc_db* db, const char* pred, orc_val* rows, orc_ann* anns) {
    int32_t ri = find_rel(db, pred);
    if (ri < 0) return ORC_ERR_UNKNOWN_PRED;
    rowset* rs = &db->rel[ri].rs;
    rs_sort(rs);
    for (uint64_t p = 0; p < rs->n; p++) {
        uint32_t id = rs->order[p];
        if (rows && rs->arity) memcpy(rows + p * rs->arity, rs_row(rs, id), rs->arity * sizeof(orc_val));
        if (anns) anns[p] = rs->anns[id];
    }
    return ORC_OK;
}

/* ------------------------------------------------------------- triangles */

/* crates/strata-gpu/src/wcoj.rs:27-43 (tri_count) over a given orientation:
 * the oriented simple graph (a[i] -> b[i], sorted by (a, b), unique) as a CSR
 * trie, and for every oriented edge (u, v) the size of out(u) ∩ out(v) by a
 * two-pointer merge.  With any acyclic orientation of a simple graph each
 * triangle is counted exactly once. */
uint64_t orc_tri_count_oriented(const uint32_t* a, const uint32_t* b, uint64_t n, uint64_t nnodes) {
    uint64_t* off = (uint64_t*)calloc(nnodes + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) off[a[i] + 1]++;
    for (uint64_t v = 0; v < nnodes; v++) off[v + 1] += off[v];
    uint64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 4096)
    for (uint64_t i = 0; i < n; i++) {
        uint32_t u = a[i], v = b[i];
        uint64_t p = off[u], pe = off[u + 1], q = off[v], qe = off[v + 1];
        uint64_t c = 0;
        while (p < pe && q < qe) {
            uint32_t x = b[p], y = b[q];
            if (x < y) p++;
            else if (x > y) q++;
            else { c++; p++; q++; }
        }
        total += c;
    }
    free(off);
    return total;
}

/* ------------------------------------------------------------------------
 * Multi-source cheapest cost by Dijkstra with a hop tie-break: an algorithm
 * independent of the semi-naive restatement above, used to check
 *   cost(S,Y) :- src(S), edge(S,Y).   cost(S,Z) :- cost(S,Y), edge(Y,Z).
 * (src facts of weight 0) at sizes the restatement cannot finish.
 *
 * cost(S,Z) is the least weight of a walk of one or more edges from S to Z,
 * and hops(S,Z) the fewest edges among the walks of that weight.  Labels are
 * ordered lexicographically by (weight, hops); with non-negative weights
 * every edge raises a label, so the first label settled for a node is its
 * least.  The heap starts from S's out-edges (one hop), not from S itself,
 * so cost(S,S) exists only through a cycle, as in the rules.
 *
 * The semi-naive round count of this stratum follows from the hops
 * (seminaive.rs:94-104): after the bootstrap the relation holds the best
 * walks of at most one edge, after delta round k those of at most k+1, and
 * round k's delta is empty exactly when no label improves any more, i.e. when
 * k reaches the largest hops(S,Z) over every source and reachable Z (0 when
 * the bootstrap derives nothing).
 * ------------------------------------------------------------------------ */

typedef struct {
    int64_t w;
    uint32_t h;
    uint32_t v;
} dj_item;

static int dj_less(const dj_item* a, const dj_item* b) {
    return a->w < b->w || (a->w == b->w && a->h < b->h);
}

typedef struct {
    dj_item* a;
    uint64_t cap;
    int oom;
} dj_heap;

static void dj_push(dj_heap* hp, uint64_t* n, dj_item x) {
    if (*n == hp->cap) {
        uint64_t cap = hp->cap ? 2 * hp->cap : 1024;
        dj_item* a = (dj_item*)realloc(hp->a, cap * sizeof(dj_item));
        if (!a) {
            hp->oom = 1;
            return;
        }
        hp->a = a;
        hp->cap = cap;
    }
    dj_item* heap = hp->a;
    uint64_t i = (*n)++;
    while (i > 0) {
        uint64_t p = (i - 1) / 2;
        if (!dj_less(&x, &heap[p])) break;
        heap[i] = heap[p];
        i = p;
    }
    heap[i] = x;
}

static dj_item dj_pop(dj_item* heap, uint64_t* n) {
    dj_item top = heap[0];
    dj_item x = heap[--(*n)];
    uint64_t i = 0;
    for (;;) {
        uint64_t c = 2 * i + 1;
        if (c >= *n) break;
        if (c + 1 < *n && dj_less(&heap[c + 1], &heap[c])) c++;
        if (!dj_less(&heap[c], &x)) break;
        heap[i] = heap[c];
        i = c;
    }
    if (*n) heap[i] = x;
    return top;
}

int32_t orc_cost_dijkstra(uint64_t nnodes, uint64_t nedges, const uint32_t* src, const uint32_t* dst,
                          const int64_t* w, uint32_t nsources, const uint32_t* sources, int64_t* out_w,
                          uint32_t* out_h, uint64_t* out_max_hops) {
    for (uint64_t i = 0; i < nedges; i++)
        if (src[i] >= nnodes || dst[i] >= nnodes || w[i] < 0) return ORC_ERR_BAD_INPUT;
    for (uint32_t s = 0; s < nsources; s++)
        if (sources[s] >= nnodes) return ORC_ERR_BAD_INPUT;
    uint64_t* off = (uint64_t*)calloc(nnodes + 1, sizeof(uint64_t));
    uint32_t* adj = (uint32_t*)malloc((nedges ? nedges : 1) * sizeof(uint32_t));
    int64_t* aw = (int64_t*)malloc((nedges ? nedges : 1) * sizeof(int64_t));
    uint64_t* fill = (uint64_t*)malloc((nnodes + 1) * sizeof(uint64_t));
    if (!off || !adj || !aw || !fill) {
        free(off); free(adj); free(aw); free(fill);
        return ORC_ERR_BAD_INPUT;
    }
    for (uint64_t i = 0; i < nedges; i++) off[src[i] + 1]++;
    for (uint64_t v = 0; v < nnodes; v++) off[v + 1] += off[v];
    memcpy(fill, off, (nnodes + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < nedges; i++) {
        uint64_t p = fill[src[i]]++;
        adj[p] = dst[i];
        aw[p] = w[i];
    }
    free(fill);
    uint64_t maxh = 0;
    int32_t rc = ORC_OK;
#pragma omp parallel reduction(max : maxh)
    {
        dj_heap hp = {NULL, 0, 0};
        uint8_t* done = (uint8_t*)malloc(nnodes ? nnodes : 1);
#pragma omp for schedule(dynamic, 1)
        for (uint32_t s = 0; s < nsources; s++) {
            int64_t* dw = out_w + (uint64_t)s * nnodes;
            uint32_t* dh = out_h + (uint64_t)s * nnodes;
            if (!done || hp.oom) {
#pragma omp atomic write
                rc = ORC_ERR_BAD_INPUT;
                continue;
            }
            for (uint64_t v = 0; v < nnodes; v++) {
                dw[v] = INT64_MAX;
                dh[v] = UINT32_MAX;
            }
            memset(done, 0, nnodes);
            uint64_t hn = 0;
            const uint32_t s0 = sources[s];
            for (uint64_t p = off[s0]; p < off[s0 + 1]; p++) {
                dj_item it = {aw[p], 1, adj[p]};
                dj_item cur = {dw[it.v], dh[it.v], it.v};
                if (dj_less(&it, &cur)) {
                    dw[it.v] = it.w;
                    dh[it.v] = 1;
                    dj_push(&hp, &hn, it);
                }
            }
            while (hn && !hp.oom) {
                dj_item x = dj_pop(hp.a, &hn);
                if (done[x.v] || x.w != dw[x.v] || x.h != dh[x.v]) continue;
                done[x.v] = 1;
                if (x.h > maxh) maxh = x.h;
                for (uint64_t p = off[x.v]; p < off[x.v + 1]; p++) {
                    const uint32_t y = adj[p];
                    if (done[y]) continue;
                    if (x.w > INT64_MAX - aw[p]) continue; /* cannot beat a finite label */
                    dj_item it = {x.w + aw[p], x.h + 1, y};
                    dj_item cur = {dw[y], dh[y], y};
                    if (dj_less(&it, &cur)) {
                        dw[y] = it.w;
                        dh[y] = it.h;
                        dj_push(&hp, &hn, it);
                    }
                }
            }
            if (hp.oom) {
#pragma omp atomic write
                rc = ORC_ERR_BAD_INPUT;
            }
        }
        free(hp.a);
        free(done);
    }
    free(off);
    free(adj);
    free(aw);
    *out_max_hops = maxh;
    return rc;
}
