/*
 * CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h for the list of
 * restated reference functions and how each part is pinned).
 *
 * Built with -ffp-contract=off so that a*b+c is never fused: the reference is
 * compiled for baseline x86-64, where no fused multiply-add is emitted, and
 * this file keeps its operation order.
 *
 * The BDF integrator's step-control constants and helper structure follow
 * SUNDIALS CVODE (cvode.c, Copyright (c) 2002-2025, Lawrence Livermore
 * National Security and Southern Methodist University; BSD-3-Clause).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ct_defs.h:84-176 */
#define AVOGADRO 6.02214076e26
#define BOLTZMANN 1.380649e-23
#define GAS_CONSTANT (AVOGADRO * BOLTZMANN)
#define SMALL_NUMBER 1e-300
#define BIG_NUMBER 1e300

#define MAXR 8   /* max distinct species per reaction side */
#define MAXTB 24 /* max explicit third-body efficiencies */

static _Thread_local char g_err[512];
const char* or_last_error(void) { return g_err; }
static void set_err(const char* s) { snprintf(g_err, sizeof g_err, "%s", s); }

/* ---- StoichManagerN restatement (StoichManager.h:131-366, 676-744) ---- */
typedef struct { int rxn, n, k[3]; } c123_t;
typedef struct { int rxn, n, k[MAXR]; double order[MAXR], stoich[MAXR]; } cn_t;
typedef struct {
    int n1, n2, n3, nn, cap;
    c123_t *c1, *c2, *c3;
    cn_t* cn;
} stoich_t;

static void stoich_init(stoich_t* s, int cap) {
    memset(s, 0, sizeof *s);
    s->cap = cap;
    s->c1 = calloc(cap, sizeof(c123_t));
    s->c2 = calloc(cap, sizeof(c123_t));
    s->c3 = calloc(cap, sizeof(c123_t));
    s->cn = calloc(cap, sizeof(cn_t));
}
static void stoich_free(stoich_t* s) { free(s->c1); free(s->c2); free(s->c3); free(s->cn); }

/* StoichManagerN::add (StoichManager.h:676-723) */
static void stoich_add(stoich_t* s, int rxn, int n, const int* k, const double* order, const double* st) {
    int frac = 0;
    for (int i = 0; i < n; i++)
        if (fmod(st[i], 1.0) != 0.0 || st[i] != order[i]) frac = 1;
    if (!frac && n <= 3) {
        int krep[8], nrep = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < st[i]; j++)
                if (nrep < 8) krep[nrep++] = k[i];
        c123_t c = {rxn, nrep, {0, 0, 0}};
        for (int j = 0; j < nrep && j < 3; j++) c.k[j] = krep[j];
        if (nrep == 1) { s->c1[s->n1++] = c; return; }
        if (nrep == 2) { s->c2[s->n2++] = c; return; }
        if (nrep == 3) { s->c3[s->n3++] = c; return; }
    }
    cn_t c;
    memset(&c, 0, sizeof c);
    c.rxn = rxn;
    c.n = n;
    for (int i = 0; i < n; i++) { c.k[i] = k[i]; c.order[i] = order[i]; c.stoich[i] = st[i]; }
    s->cn[s->nn++] = c;
}

static void stoich_multiply(const stoich_t* s, const double* S, double* R) {
    for (int i = 0; i < s->n1; i++) R[s->c1[i].rxn] *= S[s->c1[i].k[0]];
    for (int i = 0; i < s->n2; i++) {
        const c123_t* c = &s->c2[i];
        if (S[c->k[0]] < 0 && S[c->k[1]] < 0) R[c->rxn] = 0;
        else R[c->rxn] *= S[c->k[0]] * S[c->k[1]];
    }
    for (int i = 0; i < s->n3; i++) {
        const c123_t* c = &s->c3[i];
        double a = S[c->k[0]], b = S[c->k[1]], d = S[c->k[2]];
        if ((a < 0 && (b < 0 || d < 0)) || (b < 0 && d < 0)) R[c->rxn] = 0;
        else R[c->rxn] *= a * b * d;
    }
    for (int i = 0; i < s->nn; i++) {
        const cn_t* c = &s->cn[i];
        for (int j = 0; j < c->n; j++) {
            double order = c->order[j];
            if (order != 0.0) {
                double x = S[c->k[j]];
                if (x > 0.0) R[c->rxn] *= pow(x, order);
                else R[c->rxn] = 0.0;
            }
        }
    }
}

static void stoich_inc_species(const stoich_t* s, const double* R, double* S, double sign) {
    for (int i = 0; i < s->n1; i++) {
        if (sign > 0) S[s->c1[i].k[0]] += R[s->c1[i].rxn];
        else S[s->c1[i].k[0]] -= R[s->c1[i].rxn];
    }
    for (int i = 0; i < s->n2; i++) {
        const c123_t* c = &s->c2[i];
        if (sign > 0) { S[c->k[0]] += R[c->rxn]; S[c->k[1]] += R[c->rxn]; }
        else { S[c->k[0]] -= R[c->rxn]; S[c->k[1]] -= R[c->rxn]; }
    }
    for (int i = 0; i < s->n3; i++) {
        const c123_t* c = &s->c3[i];
        for (int j = 0; j < 3; j++) {
            if (sign > 0) S[c->k[j]] += R[c->rxn];
            else S[c->k[j]] -= R[c->rxn];
        }
    }
    for (int i = 0; i < s->nn; i++) {
        const cn_t* c = &s->cn[i];
        double x = R[c->rxn];
        for (int j = 0; j < c->n; j++) {
            if (sign > 0) S[c->k[j]] += c->stoich[j] * x;
            else S[c->k[j]] -= c->stoich[j] * x;
        }
    }
}

static void stoich_inc_reactions(const stoich_t* s, const double* S, double* R, double sign) {
    for (int i = 0; i < s->n1; i++) {
        if (sign > 0) R[s->c1[i].rxn] += S[s->c1[i].k[0]];
        else R[s->c1[i].rxn] -= S[s->c1[i].k[0]];
    }
    for (int i = 0; i < s->n2; i++) {
        const c123_t* c = &s->c2[i];
        double v = S[c->k[0]] + S[c->k[1]];
        if (sign > 0) R[c->rxn] += v;
        else R[c->rxn] -= v;
    }
    for (int i = 0; i < s->n3; i++) {
        const c123_t* c = &s->c3[i];
        double v = S[c->k[0]] + S[c->k[1]] + S[c->k[2]];
        if (sign > 0) R[c->rxn] += v;
        else R[c->rxn] -= v;
    }
    for (int i = 0; i < s->nn; i++) {
        const cn_t* c = &s->cn[i];
        for (int j = 0; j < c->n; j++) {
            if (sign > 0) R[c->rxn] += c->stoich[j] * S[c->k[j]];
            else R[c->rxn] -= c->stoich[j] * S[c->k[j]];
        }
    }
}

/* ---- mechanism ---- */
struct or_mech {
    int K, I, E;
    char (*names)[48];
    double *mw, *rmw, *comp, *tmin, *tmid, *tmax, *pref, *clow, *chigh;
    int *type, *rev;
    double *dn, *A, *b, *EaR, *A0, *b0, *EaR0, *ta, *rt3, *rt1, *t2;
    int *nr, *np, *ntb;
    int (*rk)[MAXR], (*pk)[MAXR], (*tbk)[MAXTB];
    double (*rs)[MAXR], (*ro)[MAXR], (*ps)[MAXR], (*tbe)[MAXTB];
    double* tbdef;
    int* pdep;   /* 4 PLOG, 5 Chebyshev (type[] is 0 for those), else 0 */
    double** pd; /* PLOG: (P, A, b, Ea/R) per entry; Chebyshev: Tmin, Tmax, Pmin, Pmax, nT, nP, coefficients */
    int* npd;    /* PLOG entries / Chebyshev record length */
    double (*fx)[5]; /* SRI a, b, c, d, e; Tsang A, B */
    int* ca;         /* chemically activated falloff */
    stoich_t reac, prod, revprod;
};

static int next_tok(char** p, char* out, int cap) {
    char* s = *p;
    while (*s == ' ' || *s == '\t') s++;
    if (!*s || *s == '\n') return 0;
    int n = 0;
    while (*s && *s != ' ' && *s != '\t' && *s != '\n') {
        if (n < cap - 1) out[n++] = *s;
        s++;
    }
    out[n] = 0;
    *p = s;
    return 1;
}
static double tokd(char** p) {
    char t[64];
    if (!next_tok(p, t, sizeof t)) return NAN;
    return strtod(t, NULL);
}
static int toki(char** p) {
    char t[64];
    if (!next_tok(p, t, sizeof t)) return -1;
    return atoi(t);
}

or_mech* or_mech_load(const char* path) {
    FILE* f = fopen(path, "r");
    if (!f) { set_err("cannot open mechanism file"); return NULL; }
    or_mech* m = calloc(1, sizeof *m);
    size_t cap = 1 << 16;
    char* line = malloc(cap);
    char tok[256];
    int ok = 0, last = -1;
    while (fgets(line, (int)cap, f)) {
        char* p = line;
        if (!next_tok(&p, tok, sizeof tok)) continue;
        if (!strcmp(tok, "elements")) {
            m->E = toki(&p);
        } else if (!strcmp(tok, "species")) {
            m->K = toki(&p);
            int K = m->K;
            m->names = calloc(K, sizeof *m->names);
            m->mw = calloc(K, 8); m->rmw = calloc(K, 8); m->comp = calloc((size_t)K * (m->E > 0 ? m->E : 1), 8);
            m->tmin = calloc(K, 8); m->tmid = calloc(K, 8); m->tmax = calloc(K, 8); m->pref = calloc(K, 8);
            m->clow = calloc(7 * K, 8); m->chigh = calloc(7 * K, 8);
        } else if (!strcmp(tok, "S")) {
            int k = toki(&p);
            next_tok(&p, m->names[k], 48);
            m->mw[k] = tokd(&p);
            m->rmw[k] = 1.0 / m->mw[k];
            m->tmin[k] = tokd(&p); m->tmid[k] = tokd(&p); m->tmax[k] = tokd(&p); m->pref[k] = tokd(&p);
            for (int j = 0; j < 7; j++) m->clow[7 * k + j] = tokd(&p);
            for (int j = 0; j < 7; j++) m->chigh[7 * k + j] = tokd(&p);
            for (int e = 0; e < m->E; e++) m->comp[k * m->E + e] = tokd(&p);
        } else if (!strcmp(tok, "reactions")) {
            int I = m->I = toki(&p);
            m->type = calloc(I, 4); m->rev = calloc(I, 4);
            m->dn = calloc(I, 8); m->A = calloc(I, 8); m->b = calloc(I, 8); m->EaR = calloc(I, 8);
            m->A0 = calloc(I, 8); m->b0 = calloc(I, 8); m->EaR0 = calloc(I, 8);
            m->ta = calloc(I, 8); m->rt3 = calloc(I, 8); m->rt1 = calloc(I, 8); m->t2 = calloc(I, 8);
            m->nr = calloc(I, 4); m->np = calloc(I, 4); m->ntb = calloc(I, 4);
            m->rk = calloc(I, sizeof *m->rk); m->pk = calloc(I, sizeof *m->pk); m->tbk = calloc(I, sizeof *m->tbk);
            m->rs = calloc(I, sizeof *m->rs); m->ro = calloc(I, sizeof *m->ro); m->ps = calloc(I, sizeof *m->ps);
            m->tbe = calloc(I, sizeof *m->tbe); m->tbdef = calloc(I, 8);
            m->pdep = calloc(I, 4); m->pd = calloc(I, sizeof *m->pd); m->npd = calloc(I, 4);
            m->fx = calloc(I, sizeof *m->fx);
            m->ca = calloc(I, 4);
            stoich_init(&m->reac, I + 1);
            stoich_init(&m->prod, I + 1);
            stoich_init(&m->revprod, I + 1);
        } else if (!strcmp(tok, "R")) {
            int i = toki(&p);
            m->type[i] = toki(&p); m->rev[i] = toki(&p); m->dn[i] = tokd(&p);
            if (m->type[i] < 0 || m->type[i] > 7) { set_err("unknown reaction type"); goto fail; }
            last = i;
            if (m->type[i] == 4 || m->type[i] == 5) { m->pdep[i] = m->type[i]; m->type[i] = 0; }
            m->A[i] = tokd(&p); m->b[i] = tokd(&p); m->EaR[i] = tokd(&p);
            m->A0[i] = tokd(&p); m->b0[i] = tokd(&p); m->EaR0[i] = tokd(&p);
            m->ta[i] = tokd(&p); m->rt3[i] = tokd(&p); m->rt1[i] = tokd(&p); m->t2[i] = tokd(&p);
            m->nr[i] = toki(&p);
            if (m->nr[i] > MAXR) { set_err("too many reactants"); goto fail; }
            for (int j = 0; j < m->nr[i]; j++) { m->rk[i][j] = toki(&p); m->rs[i][j] = tokd(&p); m->ro[i][j] = tokd(&p); }
            m->np[i] = toki(&p);
            if (m->np[i] > MAXR) { set_err("too many products"); goto fail; }
            for (int j = 0; j < m->np[i]; j++) { m->pk[i][j] = toki(&p); m->ps[i][j] = tokd(&p); }
            m->tbdef[i] = tokd(&p);
            m->ntb[i] = toki(&p);
            if (m->ntb[i] > MAXTB) { set_err("too many efficiencies"); goto fail; }
            for (int j = 0; j < m->ntb[i]; j++) { m->tbk[i][j] = toki(&p); m->tbe[i][j] = tokd(&p); }
            /* Kinetics::addReaction (Kinetics.cpp:740-781) */
            stoich_add(&m->reac, i, m->nr[i], m->rk[i], m->ro[i], m->rs[i]);
            stoich_add(&m->prod, i, m->np[i], m->pk[i], m->ps[i], m->ps[i]);
            if (m->rev[i]) stoich_add(&m->revprod, i, m->np[i], m->pk[i], m->ps[i], m->ps[i]);
        } else if (!strcmp(tok, "PLOG")) {
            if (last < 0 || m->pdep[last] != 4) { set_err("PLOG record without a PLOG reaction"); goto fail; }
            int n = toki(&p);
            if (n < 1) { set_err("PLOG reaction without rate expressions"); goto fail; }
            m->npd[last] = n;
            m->pd[last] = calloc(4 * (size_t)n, 8);
            for (int j = 0; j < 4 * n; j++) m->pd[last][j] = tokd(&p);
        } else if (!strcmp(tok, "CHEMACT")) {
            if (last < 0 || !(m->type[last] == 2 || m->type[last] == 3 || m->type[last] >= 6)) {
                set_err("CHEMACT record without a falloff reaction"); goto fail;
            }
            m->ca[last] = 1;
        } else if (!strcmp(tok, "SRI")) {
            if (last < 0 || m->type[last] != 6) { set_err("SRI record without an SRI reaction"); goto fail; }
            for (int j = 0; j < 5; j++) m->fx[last][j] = tokd(&p);
        } else if (!strcmp(tok, "TSANG")) {
            if (last < 0 || m->type[last] != 7) { set_err("TSANG record without a Tsang reaction"); goto fail; }
            for (int j = 0; j < 2; j++) m->fx[last][j] = tokd(&p);
        } else if (!strcmp(tok, "CHEB")) {
            if (last < 0 || m->pdep[last] != 5) { set_err("CHEB record without a Chebyshev reaction"); goto fail; }
            double hdr[6];
            for (int j = 0; j < 6; j++) hdr[j] = tokd(&p);
            int nT = (int)hdr[4], nP = (int)hdr[5];
            if (nT < 1 || nP < 1) { set_err("empty Chebyshev coefficient table"); goto fail; }
            m->npd[last] = 6 + nT * nP;
            m->pd[last] = calloc((size_t)m->npd[last], 8);
            memcpy(m->pd[last], hdr, sizeof hdr);
            for (int j = 0; j < nT * nP; j++) m->pd[last][6 + j] = tokd(&p);
        } else if (!strcmp(tok, "end")) {
            ok = 1;
        }
    }
    free(line);
    fclose(f);
    if (!ok || m->K <= 0) { set_err("truncated mechanism file"); or_mech_free(m); return NULL; }
    return m;
fail:
    free(line);
    fclose(f);
    or_mech_free(m);
    return NULL;
}

void or_mech_free(or_mech* m) {
    if (!m) return;
    free(m->names); free(m->mw); free(m->rmw); free(m->comp); free(m->tmin); free(m->tmid); free(m->tmax);
    free(m->pref); free(m->clow); free(m->chigh); free(m->type); free(m->rev); free(m->dn); free(m->A);
    free(m->b); free(m->EaR); free(m->A0); free(m->b0); free(m->EaR0); free(m->ta); free(m->rt3);
    free(m->rt1); free(m->t2); free(m->nr); free(m->np); free(m->ntb); free(m->rk); free(m->pk);
    free(m->tbk); free(m->rs); free(m->ro); free(m->ps); free(m->tbe); free(m->tbdef);
    if (m->pd)
        for (int i = 0; i < m->I; i++) free(m->pd[i]);
    free(m->pdep); free(m->pd); free(m->npd); free(m->fx); free(m->ca);
    if (m->reac.cap) { stoich_free(&m->reac); stoich_free(&m->prod); stoich_free(&m->revprod); }
    free(m);
}

int or_mech_nspecies(const or_mech* m) { return m->K; }
int or_mech_nreactions(const or_mech* m) { return m->I; }
int or_mech_nelements(const or_mech* m) { return m->E; }
void or_mech_mw(const or_mech* m, double* mw) { memcpy(mw, m->mw, sizeof(double) * m->K); }
void or_mech_composition(const or_mech* m, double* comp) { memcpy(comp, m->comp, sizeof(double) * m->K * m->E); }
int or_mech_species_index(const or_mech* m, const char* name) {
    for (int k = 0; k < m->K; k++)
        if (!strcmp(m->names[k], name)) return k;
    return -1;
}

/* ---- thermo (NasaPoly1.h:83-122, NasaPoly2.h:105-111) ---- */
static void nasa7(const or_mech* m, int k, const double* tt, double* cp, double* h, double* s) {
    const double* c = (tt[0] <= m->tmid[k]) ? &m->clow[7 * k] : &m->chigh[7 * k];
    double ct0 = c[0];
    double ct1 = c[1] * tt[0];
    double ct2 = c[2] * tt[1];
    double ct3 = c[3] * tt[2];
    double ct4 = c[4] * tt[3];
    *cp = ct0 + ct1 + ct2 + ct3 + ct4;
    *h = ct0 + 0.5 * ct1 + 1.0 / 3.0 * ct2 + 0.25 * ct3 + 0.2 * ct4 + c[5] * tt[4];
    *s = ct0 * tt[5] + ct1 + 0.5 * ct2 + 1.0 / 3.0 * ct3 + 0.25 * ct4 + c[6];
}
static void tpoly(double T, double* tt) {
    tt[0] = T;
    tt[1] = T * T;
    tt[2] = tt[1] * T;
    tt[3] = tt[2] * T;
    tt[4] = 1.0 / T;
    tt[5] = log(T);
}

void or_thermo(const or_mech* m, double T, double* cp_R, double* h_RT, double* s_R) {
    double tt[6];
    tpoly(T, tt);
    for (int k = 0; k < m->K; k++) nasa7(m, k, tt, &cp_R[k], &h_RT[k], &s_R[k]);
}

/* ---- workspace + state ---- */
typedef struct {
    int K, I;
    double *ym, *cp_R, *h_RT, *s_R, *grt, *conc, *concm, *delta, *rkcn, *kf, *ropf, *ropr, *ropnet, *wdot;
    double T, P, mmw, RT, dens, molar_density, pressure, ctot;
    int cv; /* constant-volume reactor: the P argument of rhs_w is the volume */
    /* rate multipliers (Kinetics::setMultiplier): mult (NULL = all 1) times,
     * for reaction mult_i only, mult_v (a transient perturbation) */
    const double* mult;
    int mult_i;
    double mult_v;
} work_t;

static work_t* work_new(const or_mech* m) {
    work_t* w = calloc(1, sizeof *w);
    int K = m->K, I = m->I;
    w->K = K; w->I = I;
    w->mult = NULL;
    w->mult_i = -1;
    w->mult_v = 1.0;
    double* buf = calloc((size_t)7 * K + 8 * (size_t)I + 8, 8);
    w->ym = buf; w->cp_R = w->ym + K; w->h_RT = w->cp_R + K; w->s_R = w->h_RT + K; w->grt = w->s_R + K;
    w->conc = w->grt + K; w->wdot = w->conc + K;
    w->concm = w->wdot + K; w->delta = w->concm + I; w->rkcn = w->delta + I; w->kf = w->rkcn + I;
    w->ropf = w->kf + I; w->ropr = w->ropf + I; w->ropnet = w->ropr + I;
    return w;
}
static void work_free(work_t* w) { if (w) { free(w->ym); free(w); } }

/* setMassFractions_NoNorm + setState_TP (Phase.cpp:362-372; ThermoPhase.cpp:187-198;
 * IdealGasPhase.h:338-354) */
static void set_state(const or_mech* m, work_t* w, double T, double P, const double* Y) {
    double sum = 0.0;
    for (int k = 0; k < m->K; k++) w->ym[k] = Y[k] * m->rmw[k];
    for (int k = 0; k < m->K; k++) sum += w->ym[k];
    w->mmw = 1.0 / sum;
    w->T = T;
    w->RT = T * GAS_CONSTANT;
    w->dens = P * w->mmw / w->RT;
    w->molar_density = w->dens / w->mmw;
    w->pressure = GAS_CONSTANT * w->molar_density * T;
    w->P = P;
}

/* PlogRate (PlogRate.cpp:92-123 setRates; PlogRate.h:110-164 updateFromStruct,
 * evalFromStruct).  The rate table is a map from ln P to a group of entries
 * [first, end) with equal pressure, plus copies of the first and last groups
 * at ln P = -1000 and +1000; the interval is found with upper_bound and each
 * end evaluates ln(1e-300 + sum of the group's A T^b exp(-Ea/RT)). */
static double plog_group(const double* d, int first, int end, double logT, double recipT) {
    double k = 1e-300;
    for (int j = first; j < end; j++) k += d[4 * j + 1] * exp(d[4 * j + 2] * logT - d[4 * j + 3] * recipT);
    return log(k);
}
static double plog_kf(const double* d, int n, double logP, double logT, double recipT) {
    /* keys[g], first[g], end[g] of the map, sentinels included */
    double keys[n + 2];
    int first[n + 2], end[n + 2], ng = 1;
    for (int j = 0; j < n; j++) {
        double lp = log(d[4 * j]);
        if (ng == 1 || keys[ng - 1] != lp) { keys[ng] = lp; first[ng] = j; end[ng] = j + 1; ng++; }
        else end[ng - 1] = j + 1;
    }
    keys[0] = -1000.0; first[0] = first[1]; end[0] = end[1];
    keys[ng] = 1000.0; first[ng] = first[ng - 1]; end[ng] = end[ng - 1]; ng++;
    int hi = 0;
    while (hi < ng && !(keys[hi] > logP)) hi++; /* upper_bound */
    if (hi == 0 || hi == ng) return NAN;        /* |ln P| >= 1000: out of range */
    int lo = hi - 1;
    double lk1 = plog_group(d, first[lo], end[lo], logT, recipT);
    double lk2 = plog_group(d, first[hi], end[hi], logT, recipT);
    double rdp = 1.0 / (keys[hi] - keys[lo]);
    return exp(lk1 + (lk2 - lk1) * (logP - keys[lo]) * rdp);
}

/* ChebyshevRate (ChebyshevRate.cpp:96-112 setLimits; ChebyshevRate.h:133-171
 * updateFromStruct, evalFromStruct): reduce every temperature row over the
 * pressure polynomials, then sum the temperature series; k = 10^logk. */
static double cheb_kf(const double* d, double log10P, double recipT) {
    const int nT = (int)d[4], nP = (int)d[5];
    const double* c = d + 6;
    double lPmin = log10(d[2]), lPmax = log10(d[3]);
    double TminInv = 1.0 / d[0], TmaxInv = 1.0 / d[1];
    double TrNum = -TminInv - TmaxInv, TrDen = 1.0 / (TmaxInv - TminInv);
    double PrNum = -lPmin - lPmax, PrDen = 1.0 / (lPmax - lPmin);
    double dot[nT];
    double Pr = (2 * log10P + PrNum) * PrDen;
    double Cnm1 = Pr, Cn = 1, Cnp1;
    for (int i = 0; i < nT; i++) dot[i] = c[i * nP];
    for (int j = 1; j < nP; j++) {
        Cnp1 = 2 * Pr * Cn - Cnm1;
        for (int i = 0; i < nT; i++) dot[i] += Cnp1 * c[i * nP + j];
        Cnm1 = Cn;
        Cn = Cnp1;
    }
    double Tr = (2 * recipT + TrNum) * TrDen;
    Cnm1 = Tr;
    Cn = 1;
    double logk = dot[0];
    for (int i = 1; i < nT; i++) {
        Cnp1 = 2 * Tr * Cn - Cnm1;
        logk += Cnp1 * dot[i];
        Cnm1 = Cn;
        Cn = Cnp1;
    }
    return pow(10, logk);
}

/* BulkKinetics::updateROP (BulkKinetics.cpp:457-545) + getNetProductionRates */
static void update_rop(const or_mech* m, work_t* w) {
    const int K = m->K, I = m->I;
    const double T = w->T, RT = w->RT;
    double tt[6];
    tpoly(T, tt);
    for (int k = 0; k < K; k++) nasa7(m, k, tt, &w->cp_R[k], &w->h_RT[k], &w->s_R[k]);
    /* getStandardChemPotentials: g0_RT*RT + RT*ln(P/Pref) */
    double tmp = log(w->pressure / m->pref[0]) * RT;
    for (int k = 0; k < K; k++) {
        double g0 = w->h_RT[k] - w->s_R[k];
        w->grt[k] = g0 * RT;
        w->grt[k] += tmp;
    }
    for (int i = 0; i < I; i++) w->delta[i] = 0.0;
    stoich_inc_reactions(&m->revprod, w->grt, w->delta, +1);
    stoich_inc_reactions(&m->reac, w->grt, w->delta, -1);
    double logStandConc = log(w->pressure / RT);
    double rrt = 1.0 / RT;
    for (int i = 0; i < I; i++) {
        if (m->rev[i]) {
            double v = exp(w->delta[i] * rrt - m->dn[i] * logStandConc);
            w->rkcn[i] = v < BIG_NUMBER ? v : BIG_NUMBER;
        } else {
            w->rkcn[i] = 0.0;
        }
    }
    /* concentrations and third bodies */
    for (int k = 0; k < K; k++) w->conc[k] = w->ym[k] * w->dens;
    w->ctot = w->dens / w->mmw;
    for (int i = 0; i < I; i++) {
        if (m->type[i] == 0) continue;
        double sum = 0.0;
        for (int j = 0; j < m->ntb[i]; j++) sum += m->tbe[i][j] * w->conc[m->tbk[i][j]];
        w->concm[i] = m->tbdef[i] * w->ctot + sum;
    }
    /* rate constants */
    double logT = log(T), recipT = 1.0 / T;
    for (int i = 0; i < I; i++) {
        if (m->pdep[i] == 4) {
            w->kf[i] = plog_kf(m->pd[i], m->npd[i], log(w->pressure), logT, recipT);
        } else if (m->pdep[i] == 5) {
            w->kf[i] = cheb_kf(m->pd[i], log10(w->pressure), recipT);
        } else if (m->type[i] <= 1) {
            w->kf[i] = m->A[i] * exp(m->b[i] * logT - m->EaR[i] * recipT);
        } else {
            /* FalloffRate::evalFromStruct (Falloff.h:154-176) with
             * updateTemp / F of Troe (Falloff.cpp:283-320), SRI (389-403),
             * Tsang (480-494) */
            double work = 0.0, work1 = 1.0;
            if (m->type[i] == 3) {
                double Fcent = (1.0 - m->ta[i]) * exp(-T * m->rt3[i]) + m->ta[i] * exp(-T * m->rt1[i]);
                if (m->t2[i] != 0.0) Fcent += exp(-m->t2[i] / T);
                work = log10(Fcent > SMALL_NUMBER ? Fcent : SMALL_NUMBER);
            } else if (m->type[i] == 7) {
                double Fcent = m->fx[i][0] + (m->fx[i][1] * T);
                work = log10(Fcent > SMALL_NUMBER ? Fcent : SMALL_NUMBER);
            } else if (m->type[i] == 6) {
                work = m->fx[i][0] * exp(-m->fx[i][1] / T);
                if (m->fx[i][2] != 0.0) work += exp(-T / m->fx[i][2]);
                work1 = m->fx[i][3] * pow(T, m->fx[i][4]);
            }
            double rc_low = m->A0[i] * exp(m->b0[i] * logT - m->EaR0[i] * recipT);
            double rc_high = m->A[i] * exp(m->b[i] * logT - m->EaR[i] * recipT);
            double pr = w->concm[i] * rc_low / (rc_high + SMALL_NUMBER);
            double F = 1.0;
            if (m->type[i] == 3 || m->type[i] == 7) {
                double lpr = log10(pr > SMALL_NUMBER ? pr : SMALL_NUMBER);
                double cc = -0.4 - 0.67 * work;
                double nn = 0.75 - 1.27 * work;
                double f1 = (lpr + cc) / (nn - 0.14 * (lpr + cc));
                double lgf = work / (1.0 + f1 * f1);
                F = pow(10.0, lgf);
            } else if (m->type[i] == 6) {
                double lpr = log10(pr > SMALL_NUMBER ? pr : SMALL_NUMBER);
                double xx = 1.0 / (1.0 + lpr * lpr);
                F = pow(work, xx) * work1;
            }
            if (m->ca[i]) {
                pr = F / (1.0 + pr);
                w->kf[i] = pr * rc_low;
            } else {
                pr *= F / (1.0 + pr);
                w->kf[i] = pr * rc_high;
            }
        }
    }
    for (int i = 0; i < I; i++) {
        double perturb = w->mult ? w->mult[i] : 1.0; /* m_perturb */
        if (i == w->mult_i) perturb *= w->mult_v;
        w->ropf[i] = w->kf[i] * perturb;
        if (m->type[i] == 1) w->ropf[i] *= w->concm[i];
        w->ropr[i] = w->ropf[i];
        w->ropr[i] *= w->rkcn[i];
    }
    stoich_multiply(&m->reac, w->conc, w->ropf);
    stoich_multiply(&m->revprod, w->conc, w->ropr);
    for (int i = 0; i < I; i++) w->ropnet[i] = w->ropf[i] - w->ropr[i];
    for (int k = 0; k < K; k++) w->wdot[k] = 0.0;
    stoich_inc_species(&m->prod, w->ropnet, w->wdot, +1);
    stoich_inc_species(&m->reac, w->ropnet, w->wdot, -1);
}

int or_rates(const or_mech* m, double T, double P, const double* Y, double* wdot,
             double* ropf, double* ropr, double* kf, double* rkcn) {
    work_t* w = work_new(m);
    set_state(m, w, T, P, Y);
    update_rop(m, w);
    if (wdot) memcpy(wdot, w->wdot, 8 * m->K);
    if (ropf) memcpy(ropf, w->ropf, 8 * m->I);
    if (ropr) memcpy(ropr, w->ropr, 8 * m->I);
    if (kf) memcpy(kf, w->kf, 8 * m->I);
    if (rkcn) memcpy(rkcn, w->rkcn, 8 * m->I);
    work_free(w);
    return 0;
}

/* BulkKinetics::getEquilibriumConstants (BulkKinetics.cpp:163-179) */
int or_equilibrium_constants(const or_mech* m, double T, double P, const double* Y, double* Kc) {
    work_t* w = work_new(m);
    set_state(m, w, T, P, Y);
    update_rop(m, w);
    for (int i = 0; i < m->I; i++) w->delta[i] = 0.0;
    stoich_inc_reactions(&m->prod, w->grt, w->delta, +1);
    stoich_inc_reactions(&m->reac, w->grt, w->delta, -1);
    double rrt = 1.0 / w->RT;
    double logStandConc = log(w->pressure / w->RT);
    for (int i = 0; i < m->I; i++) Kc[i] = exp(-w->delta[i] * rrt + m->dn[i] * logStandConc);
    work_free(w);
    return 0;
}

int or_rates_batch(const or_mech* m, int n, const double* T, const double* P, const double* Y,
                   double* wdot, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const int K = m->K;
#pragma omp parallel
    {
        work_t* w = work_new(m);
        double* y = malloc(8 * K);
#pragma omp for schedule(static)
        for (int s = 0; s < n; s++) {
            for (int k = 0; k < K; k++) y[k] = Y[(size_t)k * n + s];
            set_state(m, w, T[s], P[s], y);
            update_rop(m, w);
            for (int k = 0; k < K; k++) wdot[(size_t)k * n + s] = w->wdot[k];
        }
        free(y);
        work_free(w);
    }
    return 0;
}

/* ---- chemical equilibrium (ThermoPhase::equilibrate "HP" / "TP") ----
 * Gibbs-energy minimisation of the ideal-gas mixture at fixed elements, with
 * element potentials: the Newton iteration of Gordon & McBride, NASA RP-1311
 * (1994), eqs. 2.24-2.27 (gas phase only) with the step control of its
 * eqs. 3.1-3.3.  Unknowns per iteration: the element potentials pi_e, the
 * change of ln n (total kmol per kg) and, for HP, of ln T; species amounts
 * follow as dln n_j = -mu_j/RT + sum_e a_ej pi_e + dln n + (h_j/RT) dln T.
 * Elements absent from the mixture, and the species containing them, are
 * left out.  Iterates until every correction is below 1e-12 (relative to
 * the total for species amounts).  Returns 0, or -1 if max_iter is reached. */
static int solve_dense(int n, double* A, double* b) {
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int r = c + 1; r < n; r++)
            if (fabs(A[r * n + c]) > fabs(A[p * n + c])) p = r;
        if (A[p * n + c] == 0.0) return -1;
        if (p != c) {
            for (int k = 0; k < n; k++) { double t = A[c * n + k]; A[c * n + k] = A[p * n + k]; A[p * n + k] = t; }
            double t = b[c]; b[c] = b[p]; b[p] = t;
        }
        for (int r = c + 1; r < n; r++) {
            double f = A[r * n + c] / A[c * n + c];
            for (int k = c; k < n; k++) A[r * n + k] -= f * A[c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int c = n - 1; c >= 0; c--) {
        double v = b[c];
        for (int k = c + 1; k < n; k++) v -= A[c * n + k] * b[k];
        b[c] = v / A[c * n + c];
    }
    return 0;
}

int or_equilibrate(const or_mech* m, int hp, double T_in, double P, const double* Y_in, double* T_out, double* Y_out,
                   int max_iter, int* iters) {
    const int K = m->K, E = m->E;
    double* lnn = calloc(K, 8);
    double* mu = calloc(K, 8);
    double* hj = calloc(K, 8);
    double* cpj = calloc(K, 8);
    double* dln = calloc(K, 8);
    int* act = calloc(K, sizeof(int));
    double b0[16], bel[16];
    int eact[16], ne = 0;
    if (E > 16) return -2;
    /* element amounts [kmol/kg] and the specific enthalpy of the input state */
    double tt[6];
    tpoly(T_in, tt);
    double h0 = 0.0;
    for (int e = 0; e < E; e++) b0[e] = 0.0;
    for (int k = 0; k < K; k++) {
        double c, h, sR;
        nasa7(m, k, tt, &c, &h, &sR);
        h0 += Y_in[k] * m->rmw[k] * h * GAS_CONSTANT * T_in;
        for (int e = 0; e < E; e++) b0[e] += m->comp[k * E + e] * Y_in[k] * m->rmw[k];
    }
    double bmax = 0.0;
    for (int e = 0; e < E; e++) bmax = fmax(bmax, b0[e]);
    for (int e = 0; e < E; e++)
        if (b0[e] > 1e-14 * bmax) eact[ne++] = e;
    int nact = 0;
    for (int k = 0; k < K; k++) {
        act[k] = 1;
        for (int e = 0; e < E; e++) {
            int on = 0;
            for (int q = 0; q < ne; q++) on |= (eact[q] == e);
            if (!on && m->comp[k * E + e] != 0.0) act[k] = 0;
        }
        nact += act[k];
        lnn[k] = act[k] ? log(0.1 / K) : -INFINITY;
    }
    (void)nact;
    double lnntot = log(0.1);
    double T = hp ? 3800.0 : T_in;
    const int nu = ne + 1 + (hp ? 1 : 0);
    double A[18 * 18], x[18];
    int it, status = -1;
    const double SIZE = 18.420681;
    for (it = 0; it < max_iter; it++) {
        tpoly(T, tt);
        double ntot = exp(lnntot), sumn = 0.0;
        const double lnP = log(P / m->pref[0]);
        for (int k = 0; k < K; k++) {
            if (!act[k]) continue;
            double c, h, sR;
            nasa7(m, k, tt, &c, &h, &sR);
            hj[k] = h;
            cpj[k] = c;
            mu[k] = (h - sR) + (lnn[k] - lnntot) + lnP;
            sumn += exp(lnn[k]);
        }
        for (int i = 0; i < nu * nu; i++) A[i] = 0.0;
        for (int i = 0; i < nu; i++) x[i] = 0.0;
        for (int q = 0; q < ne; q++) bel[q] = 0.0;
        double sh = 0.0, scp = 0.0, shh = 0.0, snmu = 0.0, snhmu = 0.0;
        for (int k = 0; k < K; k++) {
            if (!act[k]) continue;
            const double nj = exp(lnn[k]);
            for (int q = 0; q < ne; q++) {
                const double aq = m->comp[k * E + eact[q]] * nj;
                if (aq == 0.0) continue;
                bel[q] += aq;
                for (int r = 0; r < ne; r++) A[q * nu + r] += aq * m->comp[k * E + eact[r]];
                x[q] += aq * mu[k];
                if (hp) A[q * nu + ne + 1] += aq * hj[k];
            }
            sh += nj * hj[k];
            scp += nj * cpj[k];
            shh += nj * hj[k] * hj[k];
            snmu += nj * mu[k];
            snhmu += nj * hj[k] * mu[k];
        }
        for (int q = 0; q < ne; q++) {
            A[q * nu + ne] = bel[q];
            A[ne * nu + q] = bel[q];
            x[q] += b0[eact[q]] - bel[q];
        }
        A[ne * nu + ne] = sumn - ntot;
        x[ne] = ntot - sumn + snmu;
        if (hp) {
            A[ne * nu + ne + 1] = sh;
            for (int q = 0; q < ne; q++) A[(ne + 1) * nu + q] = A[q * nu + ne + 1];
            A[(ne + 1) * nu + ne] = sh;
            A[(ne + 1) * nu + ne + 1] = scp + shh;
            x[ne + 1] = h0 / (GAS_CONSTANT * T) - sh + snhmu;
        }
        if (solve_dense(nu, A, x)) { status = -3; break; }
        const double dlnN = x[ne], dlnT = hp ? x[ne + 1] : 0.0;
        double big = fmax(5.0 * fabs(dlnT), 5.0 * fabs(dlnN)), lam2 = 1.0, conv = fmax(fabs(dlnT), fabs(dlnN));
        for (int k = 0; k < K; k++) {
            if (!act[k]) continue;
            double d = -mu[k] + dlnN + hj[k] * dlnT;
            for (int q = 0; q < ne; q++) d += m->comp[k * E + eact[q]] * x[q];
            dln[k] = d;
            const double lnx = lnn[k] - lnntot;
            if (lnx > -SIZE) {
                if (d > big) big = d;
                conv = fmax(conv, exp(lnn[k]) / sumn * fabs(d));
            } else if (d >= 0.0 && d - dlnN != 0.0) {
                lam2 = fmin(lam2, fabs((-lnx - 9.2103404) / (d - dlnN)));
            }
        }
        double lam = fmin(1.0, lam2);
        if (big > 2.0) lam = fmin(lam, 2.0 / big);
        for (int k = 0; k < K; k++)
            if (act[k]) lnn[k] += lam * dln[k];
        lnntot += lam * dlnN;
        if (hp) T = fmin(fmax(T * exp(lam * dlnT), 200.0), 6000.0);
        if (lam == 1.0 && conv < 1e-12) { status = 0; it++; break; }
    }
    double ysum = 0.0;
    for (int k = 0; k < K; k++) {
        Y_out[k] = act[k] ? exp(lnn[k]) * m->mw[k] : 0.0;
        ysum += Y_out[k];
    }
    for (int k = 0; k < K; k++) Y_out[k] /= ysum;
    *T_out = T;
    if (iters) *iters = it;
    free(lnn); free(mu); free(hj); free(cpj); free(dln); free(act);
    return status;
}

/* Analytic reactor Jacobians of n states (m = 1 kg), species-major inputs as
 * or_rates_batch; J[s*N*N + r*N + c] row-major per state. */
int or_jacobian_batch(const or_mech* m, int n, const double* T, const double* P, const double* Y, double* J,
                      int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const int K = m->K, N = K + 2;
    int err = 0;
#pragma omp parallel
    {
        double* y = malloc(8 * N);
#pragma omp for schedule(static)
        for (int s = 0; s < n; s++) {
            y[0] = 1.0;
            y[1] = T[s];
            for (int k = 0; k < K; k++) y[2 + k] = Y[(size_t)k * n + s];
            if (or_igcpr_jacobian(m, P[s], y, J + (size_t)s * N * N, 1)) {
#pragma omp atomic write
                err = 1;
            }
        }
        free(y);
    }
    return err ? -1 : 0;
}

/* ---- reactor RHS: IdealGasConstPressureReactor::updateState/eval and
 * ReactorNet::eval ---- */
/* IdealGasReactor::updateState + eval (src/zeroD/IdealGasReactor.cpp:38-117),
 * closed, rigid, adiabatic, V = the P argument: setState_TD(T, m/V) and
 * m c_v dT/dt = -sum_k u_k wdot_k V with u_k = (h_k/RT - 1) RT
 * (IdealGasPhase::getIntEnergy_RT) and c_v = (c_p,mole - R)/Wbar. */
static int rhs_cv(const or_mech* m, work_t* w, double V, const double* y, double* ydot) {
    const int K = m->K;
    double mass = y[0];
    double sum = 0.0;
    for (int k = 0; k < K; k++) w->ym[k] = y[2 + k] * m->rmw[k];
    for (int k = 0; k < K; k++) sum += w->ym[k];
    w->mmw = 1.0 / sum;
    w->T = y[1];
    w->RT = y[1] * GAS_CONSTANT;
    w->dens = mass / V;
    w->molar_density = w->dens / w->mmw;
    w->pressure = GAS_CONSTANT * w->molar_density * y[1];
    w->P = w->pressure;
    update_rop(m, w);
    double mcvdTdt = 0.0;
    for (int n = 0; n < K; n++) {
        double uk = (w->h_RT[n] - 1.0) * w->RT;
        mcvdTdt -= w->wdot[n] * uk * V;
        ydot[2 + n] = (w->wdot[n] * V) * m->mw[n] / mass;
    }
    double inner = 0.0;
    for (int k = 0; k < K; k++) inner += w->ym[k] * w->cp_R[k];
    double cv_mole = GAS_CONSTANT * (w->mmw * inner) - GAS_CONSTANT;
    double cv_mass = cv_mole / w->mmw;
    ydot[0] = 0.0;
    ydot[1] = mcvdTdt / (mass * cv_mass);
    for (int i = 0; i < K + 2; i++)
        if (!isfinite(ydot[i])) return 1;
    return 0;
}

static int rhs_w(const or_mech* m, work_t* w, double P, const double* y, double* ydot) {
    const int K = m->K;
    if (w->cv) return rhs_cv(m, w, P, y, ydot);
    double mass = y[0];
    set_state(m, w, y[1], P, y + 2);
    double vol = mass / w->dens;
    update_rop(m, w);
    double mcpdTdt = 0.0;
    mcpdTdt += 0.0;        /* m_Qdot */
    mcpdTdt += 0.0 * vol;  /* intrinsicHeating */
    for (int n = 0; n < K; n++) {
        double hk = w->h_RT[n] * w->RT;
        mcpdTdt -= w->wdot[n] * hk * vol;
        ydot[2 + n] = (w->wdot[n] * vol) * m->mw[n] / mass;
    }
    double inner = 0.0;
    for (int k = 0; k < K; k++) inner += w->ym[k] * w->cp_R[k];
    double cp_mole = GAS_CONSTANT * (w->mmw * inner);
    double cp_mass = cp_mole / w->mmw;
    ydot[0] = 0.0 / 1.0;
    ydot[1] = mcpdTdt / (mass * cp_mass);
    for (int i = 0; i < K + 2; i++)
        if (!isfinite(ydot[i])) return 1; /* checkFinite -> recoverable */
    return 0;
}

int or_igcpr_rhs(const or_mech* m, double P, const double* y, double* ydot) {
    work_t* w = work_new(m);
    int r = rhs_w(m, w, P, y, ydot);
    work_free(w);
    return r;
}

/* Analytic Jacobian d(ydot)/dy for y = [m, T, Y] (row-major, neq x neq).
 * Y columns: exact chain rule through C_j = rho*Y_j/W_j at fixed T (and hence
 * fixed molar density P/RT), with dwdot/dC from mass action, third-body and
 * falloff [M] dependence (rate constants and Kc at fixed T).
 * T column: one-sided difference with increment sqrt(eps)*T.
 * The GPU kernels evaluate the same closed form. */
/* d wdot_k / d C_j at fixed T into JC (K x K, row-major), concentrations
 * varying independently: mass action with the reference's stoichiometric
 * derivative rules, third-body and falloff [M] terms (d[M]/dC_j = the explicit
 * efficiency for colliders listed, else the default).  With use_default the
 * default-efficiency part, equal in every column, is accumulated per row in
 * acc[k] (add it to every column); skip_tb / skip_fo drop the [M] terms or only
 * the falloff d k/d[M] (the "skip-third-bodies" / "skip-falloff" derivative
 * settings, BulkKinetics.cpp:266-281, 627-660, 761-811). */
static void ddc_scatter(const or_mech* m, work_t* w, double* JC, double* acc, int use_default, int skip_tb,
                        int skip_fo) {
    const int K = m->K, I = m->I;
    for (int i = 0; i < I; i++) {
        const double* C = w->conc;
        double kfM = w->kf[i];
        double Mi = (m->type[i] != 0) ? w->concm[i] : 1.0;
        if (m->type[i] == 1) kfM *= Mi;
        /* products of concentrations */
        double pr_r = 1.0, pr_p = 1.0;
        for (int a = 0; a < m->nr[i]; a++) pr_r *= pow(C[m->rk[i][a]], m->ro[i][a]);
        if (m->rev[i])
            for (int a = 0; a < m->np[i]; a++) pr_p *= pow(C[m->pk[i][a]], m->ps[i][a]);
        double dR[MAXR + MAXR + MAXTB];
        int dk[MAXR + MAXR + MAXTB], nd = 0;
        for (int a = 0; a < m->nr[i]; a++) {
            double o = m->ro[i][a];
            if (o == 0.0) continue;
            double d = kfM * o * pow(C[m->rk[i][a]], o - 1.0);
            for (int c = 0; c < m->nr[i]; c++)
                if (c != a) d *= pow(C[m->rk[i][c]], m->ro[i][c]);
            dk[nd] = m->rk[i][a]; dR[nd++] = d;
        }
        if (m->rev[i]) {
            for (int a = 0; a < m->np[i]; a++) {
                double o = m->ps[i][a];
                double d = -kfM * w->rkcn[i] * o * pow(C[m->pk[i][a]], o - 1.0);
                for (int c = 0; c < m->np[i]; c++)
                    if (c != a) d *= pow(C[m->pk[i][c]], m->ps[i][c]);
                dk[nd] = m->pk[i][a]; dR[nd++] = d;
            }
        }
        if (!skip_tb && m->type[i] != 0 && (m->ntb[i] > 0 || (use_default && m->tbdef[i] != 0.0))) {
            double q; /* d ropnet / d[M] */
            double diff = pr_r - w->rkcn[i] * pr_p;
            if (m->type[i] == 1 || skip_fo) {
                /* three-body mass action; or a falloff reaction with its
                 * d k/d[M] skipped, whose rate then counts as linear in [M]
                 * (BulkKinetics::accumulateCompositionDerivatives,
                 * BulkKinetics.cpp:761-792, with skip-falloff) */
                q = w->kf[i] * diff;
            } else {
                double T = w->T, logT = log(T), recipT = 1.0 / T;
                double rc_low = m->A0[i] * exp(m->b0[i] * logT - m->EaR0[i] * recipT);
                double rc_high = m->A[i] * exp(m->b[i] * logT - m->EaR[i] * recipT);
                double Pr = Mi * rc_low / (rc_high + SMALL_NUMBER);
                double dlogF = 0.0, F = 1.0;
                if (m->type[i] == 6) {
                    /* d log10 F / d log10 Pr of SRI's F = w^(1/(1+lpr^2)) d T^e */
                    double wv = m->fx[i][0] * exp(-m->fx[i][1] / T);
                    if (m->fx[i][2] != 0.0) wv += exp(-T / m->fx[i][2]);
                    double lpr = log10(Pr > SMALL_NUMBER ? Pr : SMALL_NUMBER);
                    double u = 1.0 + lpr * lpr;
                    F = pow(wv, 1.0 / u) * m->fx[i][3] * pow(T, m->fx[i][4]);
                    dlogF = (Pr > SMALL_NUMBER) ? -log10(wv) * 2.0 * lpr / (u * u) : 0.0;
                } else if (m->type[i] == 3 || m->type[i] == 7) {
                    double Fcent;
                    if (m->type[i] == 7) {
                        Fcent = m->fx[i][0] + m->fx[i][1] * T;
                    } else {
                        Fcent = (1.0 - m->ta[i]) * exp(-T * m->rt3[i]) + m->ta[i] * exp(-T * m->rt1[i]);
                        if (m->t2[i] != 0.0) Fcent += exp(-m->t2[i] / T);
                    }
                    double lFc = log10(Fcent > SMALL_NUMBER ? Fcent : SMALL_NUMBER);
                    double lpr = log10(Pr > SMALL_NUMBER ? Pr : SMALL_NUMBER);
                    double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
                    double x = lpr + cc, den = nn - 0.14 * x;
                    double f1v = x / den;
                    double df1 = nn / (den * den);
                    F = pow(10.0, lFc / (1.0 + f1v * f1v));
                    dlogF = (Pr > SMALL_NUMBER) ? -lFc * 2.0 * f1v * df1 / ((1.0 + f1v * f1v) * (1.0 + f1v * f1v)) : 0.0;
                }
                /* dk/d[M] = k0' F/(1+Pr) * (1/(1+Pr) + dlogF) */
                double dkdM;
                if (m->ca[i]) {
                    /* k = k0 F/(1+Pr), Pr = [M] k0/kinf */
                    double dl = (Pr > SMALL_NUMBER) ? dlogF / Pr : 0.0;
                    dkdM = rc_low * (rc_low / (rc_high + SMALL_NUMBER)) * F / (1.0 + Pr) * (dl - 1.0 / (1.0 + Pr));
                } else {
                    double k0p = rc_low * rc_high / (rc_high + SMALL_NUMBER);
                    dkdM = k0p * F / (1.0 + Pr) * (1.0 / (1.0 + Pr) + dlogF);
                }
                q = dkdM * diff;
            }
            for (int a = 0; a < m->ntb[i]; a++) { dk[nd] = m->tbk[i][a]; dR[nd++] = q * m->tbe[i][a]; }
            if (use_default) {
                /* d[M]/dC_l holds the default efficiency for every l */
                for (int a = 0; a < m->nr[i]; a++) acc[m->rk[i][a]] -= m->rs[i][a] * q * m->tbdef[i];
                for (int a = 0; a < m->np[i]; a++) acc[m->pk[i][a]] += m->ps[i][a] * q * m->tbdef[i];
            }
        }
        /* scatter nu_ki * dR into JC, with nu the net stoichiometric
         * coefficient (a species on both sides, e.g. an explicit collider
         * kept in an elementary reaction, has nu = 0) */
        int nk[2 * MAXR], nnu = 0;
        double nuv[2 * MAXR];
        for (int a = 0; a < m->nr[i] + m->np[i]; a++) {
            int k = a < m->nr[i] ? m->rk[i][a] : m->pk[i][a - m->nr[i]];
            double v = a < m->nr[i] ? -m->rs[i][a] : m->ps[i][a - m->nr[i]];
            int found = 0;
            for (int c = 0; c < nnu; c++)
                if (nk[c] == k) { nuv[c] += v; found = 1; }
            if (!found) { nk[nnu] = k; nuv[nnu++] = v; }
        }
        for (int c = 0; c < nnu; c++) {
            if (nuv[c] == 0.0) continue;
            for (int d = 0; d < nd; d++) JC[nk[c] * K + dk[d]] += nuv[c] * dR[d];
        }
    }
}

static void jac_analytic(const or_mech* m, work_t* w, double P, const double* y, double* J) {
    const int K = m->K, N = K + 2;
    double* f0 = calloc(N, 8);
    double* f1 = calloc(N, 8);
    double* y1 = calloc(N, 8);
    double* JC = calloc((size_t)K * K, 8);
    double* acc = calloc(K, 8); /* constant volume: default-efficiency part of d[M]/dC, same in every column */
    memset(J, 0, 8 * (size_t)N * N);
    /* T column by difference */
    double dT = sqrt(DBL_EPSILON) * y[1];
    memcpy(y1, y, 8 * N);
    y1[1] = y[1] + dT;
    rhs_w(m, w, P, y1, f1);
    rhs_w(m, w, P, y, f0); /* leaves the base state in w */
    for (int i = 0; i < N; i++) J[i * N + 1] = (f1[i] - f0[i]) / dT;
    ddc_scatter(m, w, JC, acc, w->cv, 0, 0);
    /* chain rule to Y and the energy row */
    double rho = w->dens, Ct = w->ctot, Wbar = w->mmw;
    double inner = 0.0;
    for (int k = 0; k < K; k++) inner += w->ym[k] * w->cp_R[k];
    /* constant volume: c_v and u_k = h_k - RT in the energy row */
    double shift = w->cv ? 1.0 : 0.0;
    double cp_mass = GAS_CONSTANT * (w->mmw * inner - shift) / w->mmw;
    double H = 0.0, hD = 0.0;
    double* D = calloc(K, 8);
    double* hk = calloc(K, 8);
    for (int k = 0; k < K; k++) {
        hk[k] = (w->h_RT[k] - shift) * w->RT;
        double s = 0.0;
        for (int l = 0; l < K; l++) s += JC[k * K + l] * w->conc[l];
        D[k] = s;
        H += hk[k] * w->wdot[k];
        hD += hk[k] * D[k];
    }
    for (int j = 0; j < K; j++) {
        double hJ = 0.0;
        for (int k = 0; k < K; k++) {
            double jc = JC[k * K + j] + acc[k];
            if (w->cv) /* rho fixed: dC_l/dY_j = rho/W_j delta_lj */
                J[(2 + k) * N + 2 + j] = m->mw[k] / m->mw[j] * jc;
            else
                J[(2 + k) * N + 2 + j] = m->mw[k] / m->mw[j] * (JC[k * K + j] + (w->wdot[k] - D[k]) / Ct);
            hJ += hk[k] * jc;
        }
        if (w->cv) {
            double cvj = GAS_CONSTANT * (w->cp_R[j] - 1.0) / m->mw[j];
            J[1 * N + 2 + j] = -1.0 / (rho * cp_mass) * (rho * hJ / m->mw[j] - H * cvj / cp_mass);
        } else {
            double cpj = GAS_CONSTANT * w->cp_R[j] / m->mw[j];
            J[1 * N + 2 + j] = -1.0 / (rho * cp_mass) *
                               ((rho * hJ - Wbar * hD + Wbar * H) / m->mw[j] - H * cpj / cp_mass);
        }
    }
    free(D); free(hk); free(f0); free(f1); free(y1); free(JC); free(acc);
}

static void jac_fd(const or_mech* m, work_t* w, double P, const double* y, double* J) {
    const int N = m->K + 2;
    double* f0 = calloc(N, 8);
    double* f1 = calloc(N, 8);
    double* y1 = calloc(N, 8);
    rhs_w(m, w, P, y, f0);
    for (int j = 0; j < N; j++) {
        memcpy(y1, y, 8 * N);
        double d = fmax(fabs(y[j]), 1e-12) * sqrt(DBL_EPSILON);
        y1[j] = y[j] + d;
        rhs_w(m, w, P, y1, f1);
        for (int i = 0; i < N; i++) J[i * N + j] = (f1[i] - f0[i]) / d;
    }
    free(f0); free(f1); free(y1);
}

int or_igr_rhs(const or_mech* m, double V, const double* y, double* ydot) {
    work_t* w = work_new(m);
    w->cv = 1;
    int r = rhs_w(m, w, V, y, ydot);
    work_free(w);
    return r;
}

int or_igcpr_jacobian(const or_mech* m, double P, const double* y, double* J, int mode) {
    work_t* w = work_new(m);
    w->cv = (mode & 2) != 0;
    if (mode & 1) jac_analytic(m, w, P, y, J);
    else jac_fd(m, w, P, y, J);
    work_free(w);
    return 0;
}

/* ---- IdealGasConstPressureMoleReactor (state [T, n_1..n_K], kmol) ---- */

/* Phase::setMolesNoTruncate + setState_TP (Phase.cpp:550-568); returns the
 * mass sum_k n_k W_k (MoleReactor::setMassFromMoles, MoleReactor.cpp:43-51) */
static double set_state_moles(const or_mech* m, work_t* w, double T, double P, const double* n) {
    const int K = m->K;
    double totalMoles = 0.0, totalMass = 0.0;
    for (int k = 0; k < K; k++) totalMoles += n[k];
    for (int k = 0; k < K; k++) totalMass += n[k] * m->mw[k];
    double mmw = totalMass / totalMoles;
    for (int k = 0; k < K; k++) w->ym[k] = n[k] * (1.0 / (mmw * totalMoles));
    w->mmw = mmw;
    w->T = T;
    w->RT = T * GAS_CONSTANT;
    w->dens = P * w->mmw / w->RT;
    w->molar_density = w->dens / w->mmw;
    w->pressure = GAS_CONSTANT * w->molar_density * T;
    w->P = P;
    double mass = 0.0;
    for (int k = 0; k < K; k++) mass += n[k] * m->mw[k];
    return mass;
}

/* Phase::setMolesNoTruncate + setState_TD (IdealGasMoleReactor::updateState,
 * IdealGasMoleReactor.cpp:97-114): density mass / V, pressure from it */
static double set_state_moles_TV(const or_mech* m, work_t* w, double T, double V, const double* n) {
    double mass = set_state_moles(m, w, T, 1.0, n);
    w->dens = mass / V;
    w->molar_density = w->dens / w->mmw;
    w->pressure = GAS_CONSTANT * w->molar_density * T;
    w->P = w->pressure;
    return mass;
}

/* IdealGasConstPressureMoleReactor::updateState / eval
 * (IdealGasConstPressureMoleReactor.cpp:38-104) divided by the left-hand side
 * as ReactorNet::eval does: dT/dt = -sum_k wdot_k h_k V / (m c_p),
 * dn_k/dt = wdot_k V; state [T, n_1..n_K].  With w->cv the constant-volume
 * IdealGasMoleReactor (IdealGasMoleReactor.cpp:97-166): state [T, V, n_1..n_K],
 * dT/dt = -sum_k wdot_k u_k V / (m c_v), dV/dt = 0 (P is unused). */
static int rhs_mole(const or_mech* m, work_t* w, double P, const double* y, double* ydot) {
    const int K = m->K, cv = w->cv, s0 = 1 + cv;
    double mass, vol;
    if (cv) {
        vol = y[1];
        mass = set_state_moles_TV(m, w, y[0], vol, y + s0);
    } else {
        mass = set_state_moles(m, w, y[0], P, y + s0);
        vol = mass / w->dens;
    }
    update_rop(m, w);
    double inner = 0.0;
    for (int k = 0; k < K; k++) inner += w->ym[k] * w->cp_R[k];
    /* c_p or c_v per unit mass; h_k or u_k = h_k - RT */
    double c_mass = GAS_CONSTANT * (w->mmw * inner - cv) / w->mmw;
    double Total = mass * c_mass;
    double mcdTdt = 0.0;
    for (int k = 0; k < K; k++) {
        double hk = (w->h_RT[k] - cv) * w->RT;
        mcdTdt -= w->wdot[k] * hk * vol;
        ydot[s0 + k] = w->wdot[k] * vol;
    }
    ydot[0] = mcdTdt / Total;
    if (cv) ydot[1] = 0.0;
    for (int i = 0; i < K + s0; i++)
        if (!isfinite(ydot[i])) return 1;
    return 0;
}

/* IdealGasConstPressureMoleReactor::getJacobianElements
 * (IdealGasConstPressureMoleReactor.cpp:108-174) as a dense (K+1)^2 matrix in
 * state order: species block d wdot/dC (the -n_i/(V N) dC_i/dn_j terms left
 * out, as there), T column by a one-sided difference with increment
 * sqrt(eps) T, T row (c_p,j qdot - (m c_p) sum_k h_k dndot_k/dn_j) / (m c_p)^2. */
static void jac_mole(const or_mech* m, work_t* w, double P, const double* y, double* J, int skip_tb, int skip_fo) {
    const int cv = w->cv, s0 = 1 + cv, K = m->K, N = K + s0;
    double* f0 = calloc(N, 8);
    double* f1 = calloc(N, 8);
    double* y1 = calloc(N, 8);
    double* JC = calloc((size_t)K * K, 8);
    double* acc = calloc(K, 8);
    memset(J, 0, 8 * (size_t)N * N);
    /* netProductionRates_ddCi at the state (kf, rkcn, C, [M] of y) */
    rhs_mole(m, w, P, y, f0);
    ddc_scatter(m, w, JC, acc, 1, skip_tb, skip_fo);
    for (int k = 0; k < K; k++)
        for (int j = 0; j < K; j++) J[(s0 + k) * N + s0 + j] = JC[k * K + j] + acc[k];
    double dT = y[0] * sqrt(DBL_EPSILON);
    memcpy(y1, y, 8 * N);
    y1[0] += dT;
    rhs_mole(m, w, P, y1, f1);
    rhs_mole(m, w, P, y, f0); /* leaves the base state in w */
    for (int j = 0; j < N; j++) J[j * N + 0] = (f1[j] - f0[j]) / dT;
    double mass = 0.0;
    for (int k = 0; k < K; k++) mass += y[s0 + k] * m->mw[k];
    double vol = cv ? y[1] : mass / w->dens;
    double inner = 0.0;
    for (int k = 0; k < K; k++) inner += w->ym[k] * w->cp_R[k];
    /* m c_p and h_k (constant volume: m c_v, u_k and c_v,j = c_p,j - R) */
    double Total = mass * (GAS_CONSTANT * (w->mmw * inner - cv) / w->mmw);
    double qdot = 0.0;
    for (int k = 0; k < K; k++) qdot += ((w->h_RT[k] - cv) * w->RT) * (w->wdot[k] * vol);
    double denom = 1.0 / (Total * Total);
    for (int j = 0; j < K; j++) {
        double hs = 0.0;
        for (int k = 0; k < K; k++) hs += ((w->h_RT[k] - cv) * w->RT) * J[(s0 + k) * N + s0 + j];
        J[0 * N + s0 + j] = (GAS_CONSTANT * (w->cp_R[j] - cv) * qdot - Total * hs) * denom;
    }
    free(f0); free(f1); free(y1); free(JC); free(acc);
}

int or_mole_rhs(const or_mech* m, double P, const double* y, double* ydot, int cv) {
    work_t* w = work_new(m);
    w->cv = cv;
    int r = rhs_mole(m, w, P, y, ydot);
    work_free(w);
    return r;
}

int or_mole_jacobian(const or_mech* m, double P, const double* y, double* J, int flags) {
    work_t* w = work_new(m);
    w->cv = (flags >> 3) & 1;
    jac_mole(m, w, P, y, J, flags & 1, (flags >> 1) & 1);
    work_free(w);
    return 0;
}

/* ======================================================================
 * Variable-order BDF in Nordsieck form (CVODE method, restated).
 * ====================================================================== */
#define QMAX 5
#define LMAX (QMAX + 1)
#define MXSTEP 20000
#define MAXCOR 3
#define OR_MAXSENS 32
#define GM_MAXL 5      /* SPGMR Krylov dimension (SUNDIALS default, no restarts) */
#define EPLIN 0.05     /* CVLS linear-convergence factor */
#define MXNEF 7
#define MXNCF 10
#define MXNEF1 3
#define SMALL_NEF 2
#define LONG_WAIT 10
#define SMALL_NST 10
#define NLSCOEF 0.1
#define ETAMX1 10000.0
#define ETAMX2 10.0
#define ETAMX3 10.0
#define ETAMXF 0.2
#define ETAMIN 0.1
#define ETACF 0.25
#define ADDON 1e-6
#define BIAS1 6.0
#define BIAS2 6.0
#define BIAS3 10.0
#define THRESH 1.5
#define CRDOWN 0.3
#define DGMAX_LSETUP 0.3
#define RDIV 2.0
#define MSBP 20
#define MSBJ 51
#define DGMAX_JBAD 0.2
#define HLB_FACTOR 100.0
#define HUB_FACTOR 0.1
#define H_BIAS 0.5
#define HIN_ITERS 4
#define MIN_INC_MULT 1000.0
#define UROUND DBL_EPSILON

enum { FIRST_CALL = 0, PREV_CONV_FAIL = 1, PREV_ERR_FAIL = 2 };
enum { NO_FAILURES = 0, FAIL_BAD_J = 1, FAIL_OTHER = 2 };

struct or_reactor {
    const or_mech* m;
    work_t* w;
    int N, nnorm; /* nnorm: WRMS length (N, or N + 1 for the constant-volume V) */
    double P, rtol, atol;
    int jac_mode;
    double* zn[LMAX + 1];
    double *ewt, *acor, *y, *ftemp, *tempv, *savedJ, *M;
    int* piv;
    double tn, h, hprime, hscale, eta, etamax, hu, tau[LMAX + 2], tq[6], l[LMAX + 1];
    double rl1, gamma, gammap, gamrat, crate, delp, acnrm, saved_tq5;
    int q, qprime, qwait, L, nst, nstlp, nstlj, nscon;
    int convfail, jcur;
    long nfe, nje, nsetups, nni, ncfn, netf, nfe_dq;
    int started;
    double t_user;
    double* yout;
    /* forward sensitivities to reaction rate multipliers (staggered direct
     * method; sensitivities are not part of the step-size control) */
    int ns;
    int sidx[OR_MAXSENS];
    double* zS[OR_MAXSENS][LMAX + 1];
    double* acorS[OR_MAXSENS];
    double* sout[OR_MAXSENS];
    double rtolS, atolS;
    double *ewtS, *sp, *sm, *fsp, *fsm, *dS;
    long nfeS, nniS;
    double* mult;   /* rate multipliers (NULL until set) */
    /* IdealGasConstPressureMoleReactor with the AdaptivePreconditioner and
     * CVODES' SPGMR linear solver (state [T, n_1..n_K]) */
    int mole, skip_tb, skip_fo;
    double* atolv;                     /* per-component atol (NULL: atol) */
    double* V[GM_MAXL + 1];
    double *gvt, *gxc, *gwork;
    double Hes[GM_MAXL + 1][GM_MAXL], givens[2 * GM_MAXL], yg[GM_MAXL + 1];
    long nli, nps, njt, ncfl, nfe_ls;
};

static double wrms(const double* v, const double* w, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) { double x = v[i] * w[i]; s += x * x; }
    return sqrt(s / n);
}

static int rhs_mole(const or_mech* m, work_t* w, double P, const double* y, double* ydot);
static int rhs_r(or_reactor* r, const double* y, double* ydot) {
    return r->mole ? rhs_mole(r->m, r->w, r->P, y, ydot) : rhs_w(r->m, r->w, r->P, y, ydot);
}
static int f_eval(or_reactor* r, const double* y, double* ydot) {
    r->nfe++;
    return rhs_r(r, y, ydot);
}
static double atol_i(const or_reactor* r, int i) { return r->atolv ? r->atolv[i] : r->atol; }

static void ewt_set(or_reactor* r, const double* y) {
    for (int i = 0; i < r->N; i++) r->ewt[i] = 1.0 / (r->rtol * fabs(y[i]) + atol_i(r, i));
}

/* dense LU with partial pivoting (column-oriented, first maximal pivot) */
static int lu_factor(double* A, int n, int* piv) {
    for (int k = 0; k < n; k++) {
        int p = k;
        double amax = fabs(A[k * n + k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i * n + k]) > amax) { amax = fabs(A[i * n + k]); p = i; }
        piv[k] = p;
        if (A[p * n + k] == 0.0) return k + 1;
        if (p != k)
            for (int j = 0; j < n; j++) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
        double mult = 1.0 / A[k * n + k];
        for (int i = k + 1; i < n; i++) A[i * n + k] *= mult;
        for (int j = k + 1; j < n; j++) {
            double akj = A[k * n + j];
            if (akj != 0.0)
                for (int i = k + 1; i < n; i++) A[i * n + j] -= A[i * n + k] * akj;
        }
    }
    return 0;
}
static void lu_solve(const double* A, int n, const int* piv, double* b) {
    for (int k = 0; k < n; k++) {
        int p = piv[k];
        if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int k = 0; k < n; k++) {
        double bk = b[k];
        for (int i = k + 1; i < n; i++) b[i] -= A[i * n + k] * bk;
    }
    for (int k = n - 1; k >= 0; k--) {
        b[k] /= A[k * n + k];
        double bk = b[k];
        for (int i = 0; i < k; i++) b[i] -= A[i * n + k] * bk;
    }
}

/* Difference-quotient dense Jacobian (CVODE's DQ method) */
static int jac_dq(or_reactor* r, const double* y0, const double* fy, double* J) {
    const int N = r->N;
    double srur = sqrt(UROUND);
    double fnorm = wrms(fy, r->ewt, r->nnorm);
    double minInc = (fnorm != 0.0) ? (MIN_INC_MULT * fabs(r->h) * UROUND * N * fnorm) : 1.0;
    double* yy = calloc(N, 8);
    double* col = calloc(N, 8);
    memcpy(yy, y0, 8 * N);
    for (int j = 0; j < N; j++) {
        double ysave = yy[j];
        double inc = fmax(srur * fabs(ysave), minInc / r->ewt[j]);
        yy[j] += inc;
        r->nfe_dq++;
        int ret = rhs_w(r->m, r->w, r->P, yy, col);
        yy[j] = ysave;
        if (ret) { free(col); free(yy); return ret; }
        double inc_inv = 1.0 / inc;
        for (int i = 0; i < N; i++) J[i * N + j] = inc_inv * col[i] - inc_inv * fy[i];
    }
    free(col);
    free(yy);
    return 0;
}

/* linear-solver setup: decide on Jacobian reuse, form M = I - gamma*J, factor */
static int lsetup(or_reactor* r, const double* ypred, const double* fpred) {
    const int N = r->N;
    double dgamma = fabs(r->gamma / r->gammap - 1.0);
    int jbad = (r->nst == 0) || (r->nst >= r->nstlj + MSBJ) ||
               ((r->convfail == FAIL_BAD_J) && (dgamma < DGMAX_JBAD)) || (r->convfail == FAIL_OTHER);
    if (!jbad) {
        r->jcur = 0;
    } else {
        r->nje++;
        r->nstlj = r->nst;
        r->jcur = 1;
        if (r->mole) {
            /* ReactorNet::preconditionerSetup (ReactorNet.cpp:881-908): the
             * state clipped at the preconditioner's atol (0,
             * AdaptivePreconditioner::stateAdjustment), the reactor's
             * Jacobian elements there */
            for (int i = 0; i < N; i++) r->gwork[i] = ypred[i] > 0.0 ? ypred[i] : 0.0;
            jac_mole(r->m, r->w, r->P, r->gwork, r->savedJ, r->skip_tb, r->skip_fo);
            r->nfe_dq += 2;
        } else if (r->jac_mode == OR_JAC_ANALYTIC) {
            jac_analytic(r->m, r->w, r->P, ypred, r->savedJ);
            r->nfe_dq += 2;
        } else {
            int ret = jac_dq(r, ypred, fpred, r->savedJ);
            if (ret) return ret;
        }
    }
    for (int i = 0; i < N * N; i++) r->M[i] = -r->gamma * r->savedJ[i];
    for (int i = 0; i < N; i++) r->M[i * N + i] += 1.0;
    r->nsetups++;
    if (lu_factor(r->M, N, r->piv)) return 1;
    r->gamrat = 1.0;
    r->gammap = r->gamma;
    r->crate = 1.0;
    r->nstlp = r->nst;
    return 0;
}

/* coefficients l[] and tq[] of the fixed-leading-coefficient BDF */
static void set_bdf(or_reactor* r) {
    double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
    int q = r->q;
    r->l[0] = r->l[1] = xi_inv = xistar_inv = 1.0;
    for (int i = 2; i <= q; i++) r->l[i] = 0.0;
    alpha0 = alpha0_hat = -1.0;
    hsum = r->h;
    if (q > 1) {
        for (int j = 2; j < q; j++) {
            hsum += r->tau[j - 1];
            xi_inv = r->h / hsum;
            alpha0 -= 1.0 / j;
            for (int i = j; i >= 1; i--) r->l[i] += r->l[i - 1] * xi_inv;
        }
        alpha0 -= 1.0 / q;
        xistar_inv = -r->l[1] - alpha0;
        hsum += r->tau[q - 1];
        xi_inv = r->h / hsum;
        alpha0_hat = -r->l[1] - xi_inv;
        for (int i = q; i >= 1; i--) r->l[i] += r->l[i - 1] * xistar_inv;
    }
    /* error-test constants */
    double A1 = 1.0 - alpha0_hat + alpha0;
    double A2 = 1.0 + q * A1;
    r->tq[2] = fabs(A1 / (alpha0 * A2));
    r->tq[5] = fabs(A2 * xistar_inv / (r->l[q] * xi_inv));
    if (r->qwait == 1) {
        if (q > 1) {
            double C = xistar_inv / r->l[q];
            double A3 = alpha0 + 1.0 / q;
            double A4 = alpha0_hat + xi_inv;
            double Cpinv = (1.0 - A4 + A3) / A3;
            r->tq[1] = fabs(C * Cpinv);
        } else {
            r->tq[1] = 1.0;
        }
        hsum += r->tau[q];
        xi_inv = r->h / hsum;
        double A5 = alpha0 - 1.0 / (q + 1);
        double A6 = alpha0_hat - xi_inv;
        double Cppinv = (1.0 - A6 + A5) / A2;
        r->tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    r->tq[4] = NLSCOEF / r->tq[2];
}

/* The Nordsieck operations act on the state history zn and, identically, on
 * every sensitivity history zS[p] (CVODES keeps them in the same form). */
static void predict(or_reactor* r) {
    r->tn += r->h;
    for (int p = -1; p < r->ns; p++) {
        double** z = (p < 0) ? r->zn : r->zS[p];
        for (int k = 1; k <= r->q; k++)
            for (int j = r->q; j >= k; j--)
                for (int i = 0; i < r->N; i++) z[j - 1][i] += z[j][i];
    }
}
static void restore(or_reactor* r, double saved_t) {
    r->tn = saved_t;
    for (int p = -1; p < r->ns; p++) {
        double** z = (p < 0) ? r->zn : r->zS[p];
        for (int k = 1; k <= r->q; k++)
            for (int j = r->q; j >= k; j--)
                for (int i = 0; i < r->N; i++) z[j - 1][i] -= z[j][i];
    }
}
static void rescale(or_reactor* r) {
    for (int p = -1; p < r->ns; p++) {
        double** z = (p < 0) ? r->zn : r->zS[p];
        double factor = r->eta;
        for (int j = 1; j <= r->q; j++) {
            for (int i = 0; i < r->N; i++) z[j][i] *= factor;
            factor *= r->eta;
        }
    }
    r->h = r->hscale * r->eta;
    r->hscale = r->h;
    r->nscon = 0;
}

/* Sensitivity right-hand side J s + df/dp for parameter p at (y, s): the
 * simultaneous centred difference quotient CVODES uses without a user
 * sensitivity function (cvSensRhs1InternalDQ, CENTERED1, parameter value and
 * scale 1 as Reactor::addSensitivityReaction registers them),
 * [f(y + D s; k_p (1 + D)) - f(y - D s; k_p (1 - D))] / 2D with
 * D = min(sqrt(max(rtol, uround)), 1 / |s|_wrms) on the state's weights. */
static int sens_rhs(or_reactor* r, int p, const double* y, const double* sv, double* fS) {
    const int N = r->N;
    const double del = sqrt(fmax(r->rtol, UROUND));
    double nrm = 0.0;
    for (int i = 0; i < N; i++) nrm += (sv[i] * r->ewt[i]) * (sv[i] * r->ewt[i]);
    nrm = sqrt(nrm / r->nnorm);
    const double D = fmin(del, 1.0 / fmax(nrm, 1.0 / del));
    for (int i = 0; i < N; i++) {
        r->sp[i] = y[i] + D * sv[i];
        r->sm[i] = y[i] - D * sv[i];
    }
    r->w->mult_i = r->sidx[p];
    r->w->mult_v = 1.0 + D;
    int bad = rhs_r(r, r->sp, r->fsp);
    r->w->mult_v = 1.0 - D;
    bad |= rhs_r(r, r->sm, r->fsm);
    r->w->mult_i = -1;
    r->w->mult_v = 1.0;
    r->nfeS += 2;
    for (int i = 0; i < N; i++) fS[i] = (r->fsp[i] - r->fsm[i]) / (2.0 * D);
    return bad;
}

/* Staggered corrector for every sensitivity after the state's step passed
 * its error test: Newton on the linear BDF equation for s_n with the state's
 * iteration matrix, converged to the state's test on the sensitivity error
 * weights (rtolS |s| + atolS); at most 10 iterations. */
static int sens_correct(or_reactor* r) {
    const int N = r->N;
    for (int p = 0; p < r->ns; p++) {
        double** z = r->zS[p];
        double* acs = r->acorS[p];
        for (int i = 0; i < N; i++) {
            r->ewtS[i] = 1.0 / (r->rtolS * fabs(z[0][i]) + r->atolS);
            acs[i] = 0.0;
        }
        double crateS = 1.0, delpS = 0.0;
        for (int m = 0; m < 10; m++) {
            for (int i = 0; i < N; i++) r->dS[i] = z[0][i] + acs[i];
            if (sens_rhs(r, p, r->y, r->dS, r->tempv)) return 1;
            for (int i = 0; i < N; i++) r->dS[i] = -(r->rl1 * z[1][i] + acs[i] - r->gamma * r->tempv[i]);
            lu_solve(r->M, N, r->piv, r->dS);
            if (r->gamrat != 1.0) {
                double sc = 2.0 / (1.0 + r->gamrat);
                for (int i = 0; i < N; i++) r->dS[i] *= sc;
            }
            for (int i = 0; i < N; i++) acs[i] += r->dS[i];
            r->nniS++;
            double del = wrms(r->dS, r->ewtS, r->nnorm);
            if (m > 0) crateS = fmax(CRDOWN * crateS, del / delpS);
            if (del * fmin(1.0, crateS) / r->tq[4] <= 1.0) break;
            delpS = del;
        }
    }
    return 0;
}

static void increase_bdf(or_reactor* r) {
    double l[LMAX + 1];
    for (int i = 0; i <= QMAX; i++) l[i] = 0.0;
    double alpha1, prod, xiold, alpha0, hsum, xi;
    l[2] = alpha1 = prod = xiold = 1.0;
    alpha0 = -1.0;
    hsum = r->hscale;
    if (r->q > 1) {
        for (int j = 1; j < r->q; j++) {
            hsum += r->tau[j + 1];
            xi = hsum / r->hscale;
            prod *= xi;
            alpha0 -= 1.0 / (j + 1);
            alpha1 += 1.0 / xi;
            for (int i = j + 2; i >= 2; i--) l[i] = l[i] * xiold + l[i - 1];
            xiold = xi;
        }
    }
    double A1 = (-alpha0 - alpha1) / prod;
    int L = r->q + 1;
    for (int p = -1; p < r->ns; p++) {
        double** z = (p < 0) ? r->zn : r->zS[p];
        for (int i = 0; i < r->N; i++) z[L][i] = A1 * z[QMAX][i];
        for (int j = 2; j <= r->q; j++)
            for (int i = 0; i < r->N; i++) z[j][i] += l[j] * z[L][i];
    }
}
static void decrease_bdf(or_reactor* r) {
    double l[LMAX + 1];
    for (int i = 0; i <= QMAX; i++) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0, xi;
    for (int j = 1; j <= r->q - 2; j++) {
        hsum += r->tau[j];
        xi = hsum / r->hscale;
        for (int i = j + 2; i >= 2; i--) l[i] = l[i] * xi + l[i - 1];
    }
    for (int p = -1; p < r->ns; p++) {
        double** z = (p < 0) ? r->zn : r->zS[p];
        for (int j = 2; j < r->q; j++)
            for (int i = 0; i < r->N; i++) z[j][i] = -l[j] * z[r->q][i] + z[j][i];
    }
}
static void adjust_order(or_reactor* r, int dq) {
    if (dq == 1) increase_bdf(r);
    else if (dq == -1) decrease_bdf(r);
}

/* ---- CVLS + SPGMR for the mole reactor (SUNDIALS, restated from the
 * published algorithm: Saad & Schultz's GMRES with modified Gram-Schmidt and
 * Givens rotations, right preconditioning, scaling by the error weights) ---- */

static double dotv(const double* a, const double* b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* z = P^{-1} v with P = I - gamma_setup J the factored preconditioner */
static void psolve(or_reactor* r, const double* v, double* z) {
    if (z != v) memcpy(z, v, 8 * r->N);
    lu_solve(r->M, r->N, r->piv, z);
    r->nps++;
}

/* z = (I - gamma J) v with J v by cvLsDQJtimes' difference quotient at the
 * Newton iterate (r->y, r->ftemp): sig = 1 / |v|_wrms, up to 3 tries with sig
 * quartered after a recoverable RHS failure; 0, 1 recoverable, -1 fatal */
static int atimes(or_reactor* r, const double* v, double* z) {
    const int N = r->N;
    double sig = 1.0 / wrms(v, r->ewt, r->nnorm);
    int ret = 1;
    for (int it = 0; it < 3; it++) {
        for (int i = 0; i < N; i++) r->gwork[i] = sig * v[i] + r->y[i];
        ret = rhs_r(r, r->gwork, z);
        r->nfe_ls++;
        if (ret == 0) break;
        sig *= 0.25;
    }
    if (ret) return 1;
    double siginv = 1.0 / sig;
    for (int i = 0; i < N; i++) z[i] = siginv * (z[i] - r->ftemp[i]);
    r->njt++;
    for (int i = 0; i < N; i++) z[i] = -r->gamma * z[i] + v[i];
    return 0;
}

/* SUNModifiedGS: orthogonalise v[k] against v[0..k-1] into h[.][k-1], one
 * reorthogonalisation pass when the new norm is below 1e-3 of the old */
static double mod_gs(double** v, double (*h)[GM_MAXL], int k, int N) {
    const double FACTOR = 1000.0;
    double vk_norm = sqrt(dotv(v[k], v[k], N));
    for (int i = 0; i < k; i++) {
        h[i][k - 1] = dotv(v[i], v[k], N);
        for (int j = 0; j < N; j++) v[k][j] = -h[i][k - 1] * v[i][j] + v[k][j];
    }
    double new_norm = sqrt(dotv(v[k], v[k], N));
    double temp = FACTOR * vk_norm;
    if ((temp + new_norm) != temp) return new_norm;
    double new_norm_2 = 0.0;
    for (int i = 0; i < k; i++) {
        double np = dotv(v[i], v[k], N);
        temp = FACTOR * h[i][k - 1];
        if ((temp + np) == temp) continue;
        h[i][k - 1] += np;
        for (int j = 0; j < N; j++) v[k][j] = -np * v[i][j] + v[k][j];
        new_norm_2 += np * np;
    }
    if (new_norm_2 != 0.0) {
        double np = new_norm * new_norm - new_norm_2;
        new_norm = np > 0.0 ? sqrt(np) : 0.0;
    }
    return new_norm;
}

/* SUNQRfact, update form: rotate the new column n-1 of the Hessenberg matrix
 * by the previous rotations and add one; returns nonzero on a zero pivot */
static int qr_update(int n, double (*h)[GM_MAXL], double* q) {
    int nm1 = n - 1;
    for (int k = 0; k < nm1; k++) {
        double t1 = h[k][nm1], t2 = h[k + 1][nm1], c = q[2 * k], sn = q[2 * k + 1];
        h[k][nm1] = c * t1 - sn * t2;
        h[k + 1][nm1] = sn * t1 + c * t2;
    }
    double t1 = h[nm1][nm1], t2 = h[n][nm1], c, sn;
    if (t2 == 0.0) {
        c = 1.0; sn = 0.0;
    } else if (fabs(t2) >= fabs(t1)) {
        double t3 = t1 / t2;
        sn = -1.0 / sqrt(1.0 + t3 * t3);
        c = -sn * t3;
    } else {
        double t3 = t2 / t1;
        c = 1.0 / sqrt(1.0 + t3 * t3);
        sn = -c * t3;
    }
    q[2 * nm1] = c;
    q[2 * nm1 + 1] = sn;
    h[nm1][nm1] = c * t1 - sn * t2;
    return h[nm1][nm1] == 0.0 ? n : 0;
}

/* SUNQRsol: b <- R^{-1} Q b */
static int qr_sol(int n, double (*h)[GM_MAXL], const double* q, double* b) {
    for (int k = 0; k < n; k++) {
        double c = q[2 * k], sn = q[2 * k + 1], t1 = b[k], t2 = b[k + 1];
        b[k] = c * t1 - sn * t2;
        b[k + 1] = sn * t1 + c * t2;
    }
    for (int k = n - 1; k >= 0; k--) {
        if (h[k][k] == 0.0) return k + 1;
        b[k] /= h[k][k];
        for (int i = 0; i < k; i++) b[i] -= b[k] * h[i][k];
    }
    return 0;
}

/* cvLsSolve with SPGMR (zero initial guess, right preconditioner, both
 * scalings the error weights, tolerance eplifac tq[4] sqrt(N) in the scaled
 * 2-norm): solves (I - gamma J) x = b in place.  Returns 0 (converged, or the
 * residual reduced on the first Newton iteration), 1 recoverable, -1 fatal. */
static int gmres_solve(or_reactor* r, double* b, int mnewt) {
    const int N = r->N;
    double deltar = EPLIN * r->tq[4];
    double bnorm = wrms(b, r->ewt, r->nnorm);
    if (bnorm <= deltar) {
        if (mnewt > 0) memset(b, 0, 8 * N);
        return 0;
    }
    double delta = deltar * sqrt((double)N);
    double** V = r->V;
    for (int i = 0; i < N; i++) V[0][i] = r->ewt[i] * b[i];
    double r_norm = sqrt(dotv(V[0], V[0], N)), beta = r_norm, rho = r_norm;
    if (r_norm <= delta) {
        memset(b, 0, 8 * N);
        return 0;
    }
    memset(r->Hes, 0, sizeof r->Hes);
    double rot = 1.0;
    for (int i = 0; i < N; i++) V[0][i] *= 1.0 / r_norm;
    int krydim = 0, converged = 0;
    for (int l = 0; l < GM_MAXL; l++) {
        r->nli++;
        krydim = l + 1;
        for (int i = 0; i < N; i++) r->gvt[i] = V[l][i] / r->ewt[i];
        psolve(r, r->gvt, r->gvt);
        int ret = atimes(r, r->gvt, V[l + 1]);
        if (ret) { r->ncfl++; return ret < 0 ? -1 : 1; }
        for (int i = 0; i < N; i++) V[l + 1][i] *= r->ewt[i];
        r->Hes[l + 1][l] = mod_gs(V, r->Hes, l + 1, N);
        if (qr_update(krydim, r->Hes, r->givens)) { r->ncfl++; return 1; }
        rot *= r->givens[2 * l + 1];
        rho = fabs(rot * r_norm);
        if (rho <= delta) { converged = 1; break; }
        for (int i = 0; i < N; i++) V[l + 1][i] *= 1.0 / r->Hes[l + 1][l];
    }
    r->yg[0] = r_norm;
    for (int i = 1; i <= krydim; i++) r->yg[i] = 0.0;
    if (qr_sol(krydim, r->Hes, r->givens, r->yg)) { r->ncfl++; return 1; }
    for (int i = 0; i < N; i++) {
        double x = 0.0;
        for (int k = 0; k < krydim; k++) x += r->yg[k] * V[k][i];
        r->gxc[i] = x;
    }
    if (!converged && !(rho < beta)) { r->ncfl++; return 1; }
    for (int i = 0; i < N; i++) r->gxc[i] /= r->ewt[i];
    psolve(r, r->gxc, b);
    if (converged) return 0;
    r->ncfl++;               /* SUNLS_RES_REDUCED */
    return mnewt == 0 ? 0 : 1;
}

/* Newton iteration for the corrector; returns 0 on success, 1 recoverable failure,
 * -1 fatal */
static int nls(or_reactor* r, int nflag) {
    const int N = r->N;
    int callSetup;
    r->convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES : FAIL_OTHER;
    callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (r->nst == 0) ||
                (r->nst >= r->nstlp + MSBP) || (fabs(r->gamrat - 1.0) > DGMAX_LSETUP);
    double* ycor = r->acor;
    double* delta = r->tempv;
    int jbad = 0;
    for (;;) {
        for (int i = 0; i < N; i++) ycor[i] = 0.0;
        /* residual at the predictor */
        for (int i = 0; i < N; i++) r->y[i] = r->zn[0][i] + ycor[i];
        int ret = f_eval(r, r->y, r->ftemp);
        if (ret) return 1;
        for (int i = 0; i < N; i++) delta[i] = r->rl1 * r->zn[1][i] + ycor[i];
        for (int i = 0; i < N; i++) delta[i] = -r->gamma * r->ftemp[i] + delta[i];
        if (callSetup) {
            if (jbad) r->convfail = FAIL_BAD_J;
            ret = lsetup(r, r->zn[0], r->ftemp);
            if (ret) return ret > 0 ? 1 : -1;
            callSetup = 0;
        }
        int m = 0;
        int status;
        for (;;) {
            r->nni++;
            for (int i = 0; i < N; i++) delta[i] = -delta[i];
            if (r->mole) {
                /* iterative solve: no gamma-ratio scaling (CVLS scalesol off) */
                int ls = gmres_solve(r, delta, m);
                if (ls < 0) return -1;
                if (ls > 0) { status = 1; break; }
            } else {
                lu_solve(r->M, N, r->piv, delta);
                if (r->gamrat != 1.0) {
                    double s = 2.0 / (1.0 + r->gamrat);
                    for (int i = 0; i < N; i++) delta[i] *= s;
                }
            }
            for (int i = 0; i < N; i++) ycor[i] += delta[i];
            /* convergence test */
            double del = wrms(delta, r->ewt, r->nnorm);
            if (m > 0) r->crate = fmax(CRDOWN * r->crate, del / r->delp);
            double dcon = del * fmin(1.0, r->crate) / r->tq[4];
            if (dcon <= 1.0) {
                r->acnrm = (m == 0) ? del : wrms(ycor, r->ewt, r->nnorm);
                status = 0;
                break;
            }
            if (m >= 1 && del > RDIV * r->delp) { status = 1; break; }
            r->delp = del;
            m++;
            if (m >= MAXCOR) { status = 1; break; }
            for (int i = 0; i < N; i++) r->y[i] = r->zn[0][i] + ycor[i];
            ret = f_eval(r, r->y, r->ftemp);
            if (ret) { status = 1; break; }
            for (int i = 0; i < N; i++) delta[i] = r->rl1 * r->zn[1][i] + ycor[i];
            for (int i = 0; i < N; i++) delta[i] = -r->gamma * r->ftemp[i] + delta[i];
        }
        if (status == 0) {
            r->jcur = 0;
            for (int i = 0; i < N; i++) r->y[i] = r->zn[0][i] + ycor[i];
            return 0;
        }
        if (!r->jcur) {
            callSetup = 1;
            jbad = 1;
            continue;
        }
        return 1;
    }
}

static int do_error_test(or_reactor* r, int* nflag, double saved_t, int* nef, double* dsm) {
    *dsm = r->acnrm * r->tq[2];
    if (*dsm <= 1.0) return 0;
    (*nef)++;
    r->netf++;
    *nflag = PREV_ERR_FAIL;
    restore(r, saved_t);
    if (*nef == MXNEF) return -1;
    r->etamax = 1.0;
    if (*nef <= MXNEF1) {
        r->eta = 1.0 / (pow(BIAS2 * (*dsm), 1.0 / r->L) + ADDON);
        r->eta = fmax(ETAMIN, r->eta);
        if (*nef >= SMALL_NEF) r->eta = fmin(r->eta, ETAMXF);
        rescale(r);
        return 2;
    }
    if (r->q > 1) {
        r->eta = ETAMIN;
        adjust_order(r, -1);
        r->L = r->q;
        r->q--;
        r->qwait = r->L;
        rescale(r);
        return 2;
    }
    r->eta = ETAMIN;
    r->h *= r->eta;
    r->hscale = r->h;
    r->qwait = LONG_WAIT;
    r->nscon = 0;
    int ret = f_eval(r, r->zn[0], r->tempv);
    if (ret) return -1;
    for (int i = 0; i < r->N; i++) r->zn[1][i] = r->h * r->tempv[i];
    for (int p = 0; p < r->ns; p++) {
        if (sens_rhs(r, p, r->zn[0], r->zS[p][0], r->tempv)) return -1;
        for (int i = 0; i < r->N; i++) r->zS[p][1][i] = r->h * r->tempv[i];
    }
    return 2;
}

static void complete_step(or_reactor* r) {
    r->nst++;
    r->nscon++;
    r->hu = r->h;
    for (int i = r->q; i >= 2; i--) r->tau[i] = r->tau[i - 1];
    if (r->q == 1 && r->nst > 1) r->tau[2] = r->tau[1];
    r->tau[1] = r->h;
    for (int j = 0; j <= r->q; j++)
        for (int i = 0; i < r->N; i++) r->zn[j][i] += r->l[j] * r->acor[i];
    for (int p = 0; p < r->ns; p++)
        for (int j = 0; j <= r->q; j++)
            for (int i = 0; i < r->N; i++) r->zS[p][j][i] += r->l[j] * r->acorS[p][i];
    r->qwait--;
    if (r->qwait == 1 && r->q != QMAX) {
        memcpy(r->zn[QMAX], r->acor, 8 * r->N);
        for (int p = 0; p < r->ns; p++) memcpy(r->zS[p][QMAX], r->acorS[p], 8 * r->N);
        r->saved_tq5 = r->tq[5];
    }
}

static void set_eta(or_reactor* r) {
    if (r->eta < THRESH) {
        r->eta = 1.0;
        r->hprime = r->h;
    } else {
        r->eta = fmin(r->eta, r->etamax);
        r->hprime = r->h * r->eta;
    }
}

static void prepare_next_step(or_reactor* r, double dsm) {
    if (r->etamax == 1.0) {
        r->qwait = r->qwait > 2 ? r->qwait : 2;
        r->qprime = r->q;
        r->hprime = r->h;
        r->eta = 1.0;
        return;
    }
    double etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / r->L) + ADDON);
    if (r->qwait != 0) {
        r->eta = etaq;
        r->qprime = r->q;
        set_eta(r);
        return;
    }
    r->qwait = 2;
    double etaqm1 = 0.0, etaqp1 = 0.0;
    if (r->q > 1) {
        double ddn = wrms(r->zn[r->q], r->ewt, r->nnorm) * r->tq[1];
        etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / r->q) + ADDON);
    }
    if (r->q != QMAX && r->saved_tq5 != 0.0) {
        double cquot = (r->tq[5] / r->saved_tq5) * pow(r->h / r->tau[2], r->L);
        for (int i = 0; i < r->N; i++) r->tempv[i] = -cquot * r->zn[QMAX][i] + r->acor[i];
        double dup = wrms(r->tempv, r->ewt, r->nnorm) * r->tq[3];
        etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (r->L + 1)) + ADDON);
    }
    double etam = fmax(etaqm1, fmax(etaq, etaqp1));
    if (etam < THRESH) {
        r->eta = 1.0;
        r->qprime = r->q;
    } else if (etam == etaq) {
        r->eta = etaq;
        r->qprime = r->q;
    } else if (etam == etaqm1) {
        r->eta = etaqm1;
        r->qprime = r->q - 1;
    } else {
        r->eta = etaqp1;
        r->qprime = r->q + 1;
        memcpy(r->zn[QMAX], r->acor, 8 * r->N);
        for (int p = 0; p < r->ns; p++) memcpy(r->zS[p][QMAX], r->acorS[p], 8 * r->N);
    }
    set_eta(r);
}

/* one internal step; 0 ok, <0 failure */
static int bdf_step(or_reactor* r) {
    double saved_t = r->tn;
    int ncf = 0, nef = 0, nflag = FIRST_CALL;
    double dsm = 0.0;
    if (r->nst > 0 && r->hprime != r->h) {
        if (r->qprime != r->q) {
            adjust_order(r, r->qprime - r->q);
            r->q = r->qprime;
            r->L = r->q + 1;
            r->qwait = r->L;
        }
        rescale(r);
    }
    for (;;) {
        predict(r);
        set_bdf(r);
        r->rl1 = 1.0 / r->l[1];
        r->gamma = r->h * r->rl1;
        if (r->nst == 0) r->gammap = r->gamma;
        r->gamrat = (r->nst > 0) ? r->gamma / r->gammap : 1.0;
        int ret = nls(r, nflag);
        if (ret != 0) {
            r->ncfn++;
            restore(r, saved_t);
            if (ret < 0) return -2;
            ncf++;
            r->etamax = 1.0;
            if (ncf == MXNCF) return -3;
            r->eta = ETACF;
            nflag = PREV_CONV_FAIL;
            rescale(r);
            continue;
        }
        int e = do_error_test(r, &nflag, saved_t, &nef, &dsm);
        if (e == 2) continue;
        if (e < 0) return -4;
        break;
    }
    if (r->ns && sens_correct(r)) return -5;
    complete_step(r);
    prepare_next_step(r, dsm);
    r->etamax = (r->nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
    for (int i = 0; i < r->N; i++) r->acor[i] *= r->tq[2];
    return 0;
}

/* initial step size (CVODE's h0 estimate from an estimate of y'') */
static int initial_h(or_reactor* r, double tout) {
    const int N = r->N;
    double tdist = fabs(tout - r->tn);
    double tround = UROUND * fmax(fabs(r->tn), fabs(tout));
    if (tdist < 2.0 * tround) return -1;
    double hlb = HLB_FACTOR * tround;
    /* upper bound */
    double hub_inv = 0.0;
    for (int i = 0; i < N; i++) {
        double t1 = HUB_FACTOR * fabs(r->zn[0][i]) + (r->rtol * fabs(r->zn[0][i]) + atol_i(r, i));
        double v = fabs(r->zn[1][i]) / t1;
        if (v > hub_inv) hub_inv = v;
    }
    double hub = HUB_FACTOR * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    double hg = sqrt(hlb * hub);
    if (hub < hlb) { r->h = hg; return 0; }
    double hnew = hg;
    double* yy = calloc(N, 8);
    double* ft = calloc(N, 8);
    for (int c1 = 1; c1 <= HIN_ITERS; c1++) {
        double yddnrm = 0.0;
        int ok = 0;
        for (int c2 = 1; c2 <= HIN_ITERS; c2++) {
            for (int i = 0; i < N; i++) yy[i] = hg * r->zn[1][i] + r->zn[0][i];
            int ret = f_eval(r, yy, ft);
            if (ret == 0) {
                for (int i = 0; i < N; i++) ft[i] = (ft[i] - r->zn[1][i]) * (1.0 / hg);
                yddnrm = wrms(ft, r->ewt, r->nnorm);
                ok = 1;
                break;
            }
            hg *= 0.2;
        }
        if (!ok) { free(yy); free(ft); return -2; }
        hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
        if (c1 == HIN_ITERS) break;
        double hrat = hnew / hg;
        if (hrat > 0.5 && hrat < 2.0) break;
        if (c1 > 1 && hrat > 2.0) { hnew = hg; break; }
        hg = hnew;
    }
    free(yy);
    free(ft);
    double h0 = H_BIAS * hnew;
    if (h0 < hlb) h0 = hlb;
    if (h0 > hub) h0 = hub;
    r->h = h0;
    return 0;
}

or_reactor* or_reactor_new(const or_mech* m, double T0, double P, const double* Y0,
                           double rtol, double atol, int jac_mode) {
    or_reactor* r = calloc(1, sizeof *r);
    r->m = m;
    r->w = work_new(m);
    r->N = m->K + 2;
    r->P = P;
    r->rtol = rtol;
    r->atol = atol;
    /* bit 1 of the mode: IdealGasReactor (V = 1 m^3; r->P then holds V) */
    r->w->cv = (jac_mode & 2) != 0;
    r->nnorm = r->N + r->w->cv;
    r->jac_mode = jac_mode & 1;
    int N = r->N;
    for (int j = 0; j <= LMAX; j++) r->zn[j] = calloc(N, 8);
    r->ewt = calloc(N, 8); r->acor = calloc(N, 8); r->y = calloc(N, 8); r->ftemp = calloc(N, 8);
    r->tempv = calloc(N, 8); r->savedJ = calloc((size_t)N * N, 8); r->M = calloc((size_t)N * N, 8);
    r->piv = calloc(N, sizeof(int));
    r->yout = calloc(N, 8);
    /* IdealGasConstPressureReactor / IdealGasReactor::getState with V = 1 m^3 */
    set_state(m, r->w, T0, P, Y0);
    r->zn[0][0] = r->w->dens * 1.0;
    if (r->w->cv) r->P = 1.0;
    r->zn[0][1] = T0;
    memcpy(r->zn[0] + 2, Y0, 8 * m->K);
    memcpy(r->yout, r->zn[0], 8 * N);
    r->q = 1;
    r->L = 2;
    r->qwait = r->L;
    r->etamax = ETAMX1;
    r->tn = 0.0;
    r->t_user = 0.0;
    r->rtolS = 1.0e-4;   /* ReactorNet defaults (ReactorNet.h:484-486) */
    r->atolS = 1.0e-6;
    return r;
}

void or_reactor_free(or_reactor* r) {
    if (!r) return;
    for (int j = 0; j <= LMAX; j++) free(r->zn[j]);
    free(r->ewt); free(r->acor); free(r->y); free(r->ftemp); free(r->tempv); free(r->savedJ);
    free(r->M); free(r->piv); free(r->yout);
    for (int p = 0; p < r->ns; p++) {
        for (int j = 0; j <= LMAX; j++) free(r->zS[p][j]);
        free(r->acorS[p]); free(r->sout[p]);
    }
    free(r->ewtS); free(r->sp); free(r->sm); free(r->fsp); free(r->fsm); free(r->dS); free(r->mult);
    free(r->atolv);
    for (int j = 0; j <= GM_MAXL; j++) free(r->V[j]);
    free(r->gvt); free(r->gxc); free(r->gwork);
    work_free(r->w);
    free(r);
}

static int start(or_reactor* r, double tout) {
    ewt_set(r, r->zn[0]);
    if (f_eval(r, r->zn[0], r->zn[1])) return -1;
    if (initial_h(r, tout)) return -1;
    for (int i = 0; i < r->N; i++) r->zn[1][i] *= r->h;
    for (int p = 0; p < r->ns; p++) {
        /* s(0) = 0; zS[1] = h (J 0 + df/dp) */
        for (int i = 0; i < r->N; i++) r->zS[p][0][i] = 0.0;
        ewt_set(r, r->zn[0]);
        if (sens_rhs(r, p, r->zn[0], r->zS[p][0], r->zS[p][1])) return -1;
        for (int i = 0; i < r->N; i++) r->zS[p][1][i] *= r->h;
    }
    r->hscale = r->h;
    r->hprime = r->h;
    r->started = 1;
    return 0;
}

/* one CVode(ONE_STEP) call */
static int one_step(or_reactor* r) {
    if (r->nst > 0) ewt_set(r, r->zn[0]);
    double nrm = wrms(r->zn[0], r->ewt, r->nnorm);
    if (UROUND * nrm > 1.0) return -10;
    int ret = bdf_step(r);
    return ret;
}

int or_reactor_advance(or_reactor* r, double tout) {
    if (tout == r->t_user) return 0;
    if (tout < r->t_user) return -20;
    if (!r->started && start(r, tout)) return -21;
    int nsteps = 0;
    while (r->tn < tout) {
        if (nsteps >= MXSTEP) return -22;
        int ret = one_step(r);
        if (ret) return ret;
        nsteps++;
    }
    /* CVodeGetDky(tout, 0), CVodeGetSensDky(tout, 0) */
    double s = (tout - r->tn) / r->h;
    for (int p = -1; p < r->ns; p++) {
        double** z = (p < 0) ? r->zn : r->zS[p];
        double* out = (p < 0) ? r->yout : r->sout[p];
        for (int i = 0; i < r->N; i++) {
            double v = z[r->q][i];
            for (int j = r->q - 1; j >= 0; j--) v = z[j][i] + s * v;
            out[i] = v;
        }
    }
    r->t_user = tout;
    return 0;
}

/* IdealGasConstPressureMoleReactor in a ReactorNet with an
 * AdaptivePreconditioner and the GMRES linear solver
 * (IdealGasConstPressureMoleReactor.cpp; ReactorNet.cpp:330-339,881-908;
 * CVodesIntegrator.cpp:424-434): V = 1 m^3 at (T0, P, Y0), state
 * [T, n_1..n_K].  flags: 1 skip third-body derivatives, 2 skip falloff
 * derivatives, 4 atol as given for every component (ReactorNet::setTolerances);
 * otherwise the species' atol is atol times the initial total moles
 * (MoleReactor::updateDefaultTolerances, MoleReactor.cpp:64-76); 8 the
 * constant-volume IdealGasMoleReactor, state [T, V, n_1..n_K]. */
or_reactor* or_mole_reactor_new(const or_mech* m, double T0, double P, const double* Y0, double rtol, double atol,
                                int flags) {
    or_reactor* r = or_reactor_new(m, T0, P, Y0, rtol, atol, 0);
    const int cv = (flags >> 3) & 1, s0 = 1 + cv, K = m->K, N = K + s0;
    r->mole = 1;
    r->skip_tb = flags & 1;
    r->skip_fo = (flags >> 1) & 1;
    r->w->cv = cv;
    r->N = N;
    r->nnorm = N;
    set_state(m, r->w, T0, P, Y0);
    double mass = r->w->dens * 1.0;
    r->zn[0][0] = T0;
    if (cv) r->zn[0][1] = 1.0;   /* V = 1 m^3 */
    double tot = 0.0;
    for (int k = 0; k < K; k++) {
        r->zn[0][s0 + k] = mass * m->rmw[k] * Y0[k];
        tot += r->zn[0][s0 + k];
    }
    memcpy(r->yout, r->zn[0], 8 * N);
    r->atolv = calloc(N, 8);
    for (int i = 0; i < N; i++) r->atolv[i] = (i >= s0 && !(flags & 4)) ? atol * tot : atol;
    for (int j = 0; j <= GM_MAXL; j++) r->V[j] = calloc(N, 8);
    r->gvt = calloc(N, 8); r->gxc = calloc(N, 8); r->gwork = calloc(N, 8);
    return r;
}

/* CVodesIntegrator::solverStats linear-solver fields (CVodesIntegrator.cpp:
 * 572-621): lin_iters, prec_evals, prec_solves, jt_vec_prod_evals,
 * lin_rhs_evals, lin_conv_fails */
void or_reactor_lin_stats(const or_reactor* r, long* s) {
    s[0] = r->nli; s[1] = r->nje; s[2] = r->nps; s[3] = r->njt; s[4] = r->nfe_ls; s[5] = r->ncfl;
}

/* Reactor::addSensitivityReaction (src/zeroD/Reactor.cpp:278-289): parameter
 * = a multiplier of reaction rxn's rate (nominal 1), before the first step. */
int or_reactor_add_sensitivity(or_reactor* r, int rxn) {
    if (r->started) return -1;
    if (rxn < 0 || rxn >= r->m->I || r->ns >= OR_MAXSENS) return -2;
    const int N = r->N, p = r->ns;
    if (!r->ewtS) {
        r->ewtS = calloc(N, 8); r->sp = calloc(N, 8); r->sm = calloc(N, 8);
        r->fsp = calloc(N, 8); r->fsm = calloc(N, 8); r->dS = calloc(N, 8);
    }
    for (int j = 0; j <= LMAX; j++) r->zS[p][j] = calloc(N, 8);
    r->acorS[p] = calloc(N, 8);
    r->sout[p] = calloc(N, 8);
    r->sidx[p] = rxn;
    r->ns++;
    return 0;
}
void or_reactor_set_sensitivity_tolerances(or_reactor* r, double rtol, double atol) {
    r->rtolS = rtol;
    r->atolS = atol;
}
/* dy_i/dp at the output time, S[i * ns + p] (CVodesIntegrator::sensitivity,
 * src/numerics/CVodesIntegrator.cpp:623-644: zero before the first step). */
void or_reactor_sensitivities(const or_reactor* r, double* S) {
    for (int i = 0; i < r->N; i++)
        for (int p = 0; p < r->ns; p++) S[i * r->ns + p] = r->started ? r->sout[p][i] : 0.0;
}
/* Kinetics::setMultiplier for a whole run (all reactions at once). */
void or_reactor_set_multipliers(or_reactor* r, const double* mult) {
    if (!r->mult) r->mult = calloc(r->m->I, 8);
    memcpy(r->mult, mult, 8 * r->m->I);
    r->w->mult = r->mult;
}

int or_reactor_step(or_reactor* r) {
    if (!r->started && start(r, r->t_user + 1.0)) return -21;
    int ret = one_step(r);
    if (ret) return ret;
    memcpy(r->yout, r->zn[0], 8 * r->N);
    for (int p = 0; p < r->ns; p++) memcpy(r->sout[p], r->zS[p][0], 8 * r->N);
    r->t_user = r->tn;
    return 0;
}

double or_reactor_time(const or_reactor* r) { return r->t_user; }
void or_reactor_state(const or_reactor* r, double* y) { memcpy(y, r->yout, 8 * r->N); }
void or_reactor_stats(const or_reactor* r, long* s) {
    s[0] = r->nst; s[1] = r->nfe; s[2] = r->nje; s[3] = r->nsetups; s[4] = r->nni;
    s[5] = r->ncfn; s[6] = r->netf; s[7] = r->q; s[8] = r->nfe_dq;
}

int or_reactors_run(const or_mech* m, int n, const double* T0, const double* P, const double* Y0,
                    const double* dt, int nout, double rtol, double atol, int jac_mode,
                    double* y_out, long* stats_out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const int K = m->K, N = K + 2;
    int fail = 0;
#pragma omp parallel for schedule(static, 1)
    for (int i = 0; i < n; i++) {
        or_reactor* r = or_reactor_new(m, T0[i], P[i], Y0 + (size_t)i * K, rtol, atol, jac_mode);
        if (y_out) or_reactor_state(r, y_out + ((size_t)i * (nout + 1)) * N);
        for (int k = 1; k <= nout; k++) {
            int ret = or_reactor_advance(r, k * dt[i]);
            if (ret) {
#pragma omp critical
                { if (!fail) fail = -(i + 1); }
                break;
            }
            if (y_out) or_reactor_state(r, y_out + ((size_t)i * (nout + 1) + k) * N);
        }
        if (stats_out) or_reactor_stats(r, stats_out + (size_t)i * 9);
        or_reactor_free(r);
    }
    return fail;
}

/* Batch of mole reactors (IdealGasConstPressureMoleReactor + adaptive
 * preconditioner + GMRES), OpenMP over reactors as or_reactors_run; the final
 * states [T, n_1..n_K] into y_out (n x (K+1)) when given. */
int or_mole_reactors_run(const or_mech* m, int n, const double* T0, const double* P, const double* Y0,
                         const double* dt, int nout, double rtol, double atol, int flags, double* y_out,
                         int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const int K = m->K;
    int fail = 0;
#pragma omp parallel for schedule(static, 1)
    for (int i = 0; i < n; i++) {
        or_reactor* r = or_mole_reactor_new(m, T0[i], P[i], Y0 + (size_t)i * K, rtol, atol, flags);
        for (int k = 1; k <= nout; k++) {
            if (or_reactor_advance(r, k * dt[i])) {
#pragma omp critical
                { if (!fail) fail = -(i + 1); }
                break;
            }
        }
        if (y_out) or_reactor_state(r, y_out + (size_t)i * r->N);
        or_reactor_free(r);
    }
    return fail;
}
