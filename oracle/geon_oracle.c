/*
 * geon_oracle.c — FP64 CPU restatement of Geon's per-substep physics core.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP backend is compared against,
 * and the CPU baseline timed by bench.py.  The product library never links it.
 *
 * Restates (a9lim/physsim):
 *   src/forces.js      pairForce 163-442, calculateForce 622-694,
 *                      computeAllForces 61-154, _accum1PN 445-506,
 *                      _compute1PNTreeWalk 511-571, compute1PN 578-610
 *   src/quadtree.js    pool 13-103, _subdivide 119-172, _childFor 175-179,
 *                      insert 185-237, calculateMassDistribution 239-312,
 *                      query 372-404, build 412-419
 *   src/integrator.js  _syncToggles 138-150, _addGhost/_generateGhosts 178-244,
 *                      _buildTree 247-260, external fields 553-592,
 *                      update 726-1825 (radiation, bosons, collisions-merge,
 *                      Hawking and disintegration left out), _retireParticle
 *                      152-173, history sampling 1549-1565
 *   src/signal-delay.js getDelayedState 16-221
 *   main.js            dead-particle GC 874-885
 *   src/topology.js    minImage 18-76, wrapPosition 81-135
 *   src/relativity.js  angwToAngVel 7-9, setVelocity 15-25
 *   src/energy.js      computeEnergies 11-195 (passes 1-3)
 *   src/collisions.js  overlap detection 35-60
 *   main.js            addParticle normalization 504-539
 *
 * Parity pins for this file live in tests/test_oracle.py (closed forms,
 * invariants, brute force).  Signal delay follows GeonParams.signal_delay:
 * FOLLOW couples it to GEON_RELATIVITY as the reference does (forces.js:85);
 * OFF gives current-time leaf sources.
 */
#include "geon_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants: src/config.js ---- */
#define EPS_ 1e-12                 /* EPSILON (config.js:30) */
#define INERTIA_K 0.4              /* config.js:52 */
#define MAG_MOMENT_K 0.2           /* config.js:53 */
#define TIDAL_STRENGTH 64.0        /* config.js:54 */
#define HIGGS_MASS_FLOOR 0.05      /* config.js:90 */
#define MAX_SPEED_RATIO 0.999      /* config.js:32 */
#define DESPAWN_MARGIN 64.0        /* config.js:138 */
#define BOSON_CHARGE 0.1           /* config.js:75 */
#define QT_CAP 4                   /* QUADTREE_CAPACITY (config.js:59) */
#define SCALAR_FIELD_MAX 2.0       /* config.js:85 */
#define HIGGS_COUPLING 1.0         /* config.js:89 */
#define HIGGS_MASS_MAX_DELTA 8.0   /* config.js:91 */
#define SELFGRAV_PHI_MAX 0.25      /* config.js:92 */
#define AXION_COUPLING 0.05        /* config.js:96 */
#define HIGGS_AXION_COUPLING 0.01  /* config.js:97 */
#define HIST_N 256                 /* HISTORY_SIZE (config.js:120) */
#define HIST_MASK (HIST_N - 1)     /* HISTORY_MASK (config.js:121) */

/* Signal-delay history ring of one particle (particle.js:77-85, 124-134). */
typedef struct {
    double x[HIST_N], y[HIST_N], vx[HIST_N], vy[HIST_N], aw[HIST_N], t[HIST_N];
    int head, count;
} Hist;

/* One scalar field on a G x G grid (scalar-field.js:13-64). */
typedef struct {
    int G;
    double mass, vac;
    int higgs;
    double *field, *fieldDot, *source, *thermal, *rho, *gradX, *gradY;
    double *sgPhi, *sgGx, *sgGy, *fre, *fim, *gre, *gim;
    double gW, gH, gSoft2; int gPer, gTopo, gValid;
    int hasEnergy;
} OField;

typedef struct {
    double x, y, vx, vy, wx, wy;
    double mass, baseMass, charge, angw, angVel;
    double radius, radiusSq, bodyRadiusSq, invMass;
    double magMoment, angMomentum;
    double axMod, yukMod, higgsMod;
    double fx, fy;
    double Bz, dBzdx, dBzdy, Bgz, dBgzdx, dBgzdy;
    double frameDrag, tidal, contactTorque;
    double f1x, f1y, f1ox, f1oy;
    double jx, jy;                 /* analytic jerk dF/dt (radiation on, forces.js:198-402) */
    double radAccum, quadAccum, emQuadAccum;   /* _radAccum, _quadAccum, _emQuadAccum */
    double pe;                     /* this particle's share of the PE accumulator */
    int32_t id;                    /* upload index */
    int32_t q_quanta;              /* round(charge / 0.1) */
    int32_t isGhost, original;     /* ghost: index of the real particle */
    int32_t antimatter;            /* particle.js:21 */
    Hist* hist;                    /* NULL until the first sample (particle.js:124) */
    double creationTime, deathTime, deathMass;   /* particle.js:17-18; _deathMass */
} P;

typedef struct {
    int cap, count;
    double *bx, *by, *bw, *bh;
    double *mass, *charge, *magMom, *angMom, *momX, *momY, *comX, *comY;
    double *mom;                   /* 9 per node: mass quadrupole, charge dipole, charge quadrupole (about the COM), sum |q| */
    int32_t *nw, *ne, *sw, *se, *pointCount, *divided, *level, *subCount, *qq;
    uint64_t *path;
    int32_t *points;               /* cap * QT_CAP indices into the tree particle array */
} Pool;

struct OrcSys {
    GeonParams prm;
    int n, capacity;
    P* p;                          /* real particles, array order */
    P* tp;                         /* tree particles: real then ghosts */
    int ntp, tpcap;
    Pool pool;
    int root;
    int forcesInit;
    double simTime;
    double potentialEnergy;
    int64_t cnt_pair, cnt_node, cnt_visit;
    /* derived toggles (integrator.js:138-150) */
    int grav, coul, mag, gm, onePN, rel, spinOrbit, bh, yuk, higgs, axion, barnesHut, expansion, rad;
    double totalRadiated, totalRadiatedPx, totalRadiatedPy;   /* sim.totalRadiated* (main.js) */
    double softeningSq, softening, yukCutoffSq;
    int needAxMod, periodic;
    OField higgsF, axionF;
    double* farModel;              /* learned far-field model (FAR_MODEL doubles) or NULL */
    int delay;                     /* signal delay in effect */
    int delay1pn;                  /* compute1PN takes leaf sources from the light cone */
    int histStride;                /* _histStride (integrator.js:111) */
    P* dead;                       /* sim.deadParticles (main.js:103) */
    int ndead, deadcap;
    int32_t next_id;               /* Particle.nextId (particle.js:13) */
    double* ev;                    /* collision events of the last update, ORC_EVENT_STRIDE doubles each */
    int nev, evcap;
};

static int g_threads = 1;
void orc_set_threads(int32_t n) { g_threads = n > 0 ? n : 1; }

/* ------------------------------------------------------------------ */
/* topology.js                                                         */
/* ------------------------------------------------------------------ */
static inline double torusWrap(double d, double full, double half) {      /* topology.js:9-12 */
    if (d > half) d -= full; else if (d < -half) d += full;
    return d;
}

static void minImage(double ox, double oy, double sx, double sy, int topo,
                     double W, double H, double halfW, double halfH,
                     double* outx, double* outy) {                        /* topology.js:18-76 */
    double dx = sx - ox, dy = sy - oy;
    if (topo == GEON_TORUS) {
        *outx = torusWrap(dx, W, halfW);
        *outy = torusWrap(dy, H, halfH);
        return;
    }
    double doubleH = 2 * H;
    if (topo == GEON_KLEIN) {
        double bx = torusWrap(dx, W, halfW), by = dy;
        double bestDx = bx, bestDy = by, bestSq = bx * bx + by * by;
        double cx = torusWrap((W - sx) - ox, W, halfW);
        double cy = torusWrap((sy + H) - oy, doubleH, H);
        double sq = cx * cx + cy * cy;
        if (sq < bestSq) { bestDx = cx; bestDy = cy; }
        *outx = bestDx; *outy = bestDy;
        return;
    }
    double doubleW = 2 * W;
    double bestDx = dx, bestDy = dy, bestSq = dx * dx + dy * dy;
    double cx = (W - sx) - ox, cy = torusWrap((sy + H) - oy, doubleH, H);
    double sq = cx * cx + cy * cy;
    if (sq < bestSq) { bestDx = cx; bestDy = cy; bestSq = sq; }
    cx = torusWrap((sx + W) - ox, doubleW, W); cy = (H - sy) - oy;
    sq = cx * cx + cy * cy;
    if (sq < bestSq) { bestDx = cx; bestDy = cy; bestSq = sq; }
    cx = torusWrap((2 * W - sx) - ox, doubleW, W); cy = torusWrap((2 * H - sy) - oy, doubleH, H);
    sq = cx * cx + cy * cy;
    if (sq < bestSq) { bestDx = cx; bestDy = cy; }
    *outx = bestDx; *outy = bestDy;
}

static void wrapPosition(P* p, int topo, double W, double H) {            /* topology.js:81-135 */
    if (topo == GEON_TORUS) {
        if (p->x < 0) p->x += W; else if (p->x > W) p->x -= W;
        if (p->y < 0) p->y += H; else if (p->y > H) p->y -= H;
        return;
    }
    if (topo == GEON_KLEIN) {
        if (p->x < 0) p->x += W; else if (p->x > W) p->x -= W;
        if (p->y < 0 || p->y > H) {
            p->y += (p->y < 0) ? H : -H;
            p->x = W - p->x;
            p->wx = -p->wx; p->vx = -p->vx; p->angw = -p->angw; p->angVel = -p->angVel;
        }
        return;
    }
    for (int iter = 0; iter < 2; iter++) {
        int wrapped = 0;
        if (p->x < 0 || p->x > W) {
            p->x += (p->x < 0) ? W : -W; p->y = H - p->y;
            p->wy = -p->wy; p->vy = -p->vy; p->angw = -p->angw; p->angVel = -p->angVel;
            wrapped = 1;
        }
        if (p->y < 0 || p->y > H) {
            p->y += (p->y < 0) ? H : -H; p->x = W - p->x;
            p->wx = -p->wx; p->vx = -p->vx; p->angw = -p->angw; p->angVel = -p->angVel;
            wrapped = 1;
        }
        if (!wrapped) break;
    }
}

/* ------------------------------------------------------------------ */
/* relativity.js                                                       */
/* ------------------------------------------------------------------ */
static inline double angwToAngVel(double angw, double radius) {           /* relativity.js:7-9 */
    return angw / sqrt(1 + angw * angw * radius * radius);
}

/* ------------------------------------------------------------------ */
/* forces.js: pairForce                                                */
/* ------------------------------------------------------------------ */
/* Observer p gets the force of a source.  Restates forces.js:163-442; with
 * radiation on it also accumulates the analytic jerk dF/dt of each term
 * (forces.js:198-402), the input of the radiation reaction. */
static void pairForce(const OrcSys* s, P* p, double sx, double sy, double svx, double svy,
                      double sMass, double sCharge, double sAngVel, double sMagMoment,
                      double sAngMomentum, double sAxMod, double sYukMod, double sHiggsMod,
                      int accumulatePE, int signalDelayed) {
    const GeonParams* g = &s->prm;
    double rx, ry;
    if (s->periodic) {
        minImage(p->x, p->y, sx, sy, g->topology, g->domain_w, g->domain_h,
                 g->domain_w * 0.5, g->domain_h * 0.5, &rx, &ry);
    } else {
        rx = sx - p->x; ry = sy - p->y;
    }
    const double rawRSq = rx * rx + ry * ry;
    const double rSq = rawRSq + s->softeningSq;
    const double invRSq = 1 / rSq;
    const double invR = sqrt(invRSq);
    const double invR3 = invR * invRSq;
    const double invR5 = invR3 * invRSq;
    double aberr = 1;                                                     /* forces.js:182-189 */
    if (signalDelayed) {
        const double nDotV = -(rx * svx + ry * svy) * invR;
        const double denom = fmax(1 - nDotV, 0.01);
        aberr = fmin(1 / (denom * denom * denom), 100);
    }
    const double invR3a = signalDelayed ? invR3 * aberr : invR3;
    const double invR5a = signalDelayed ? invR5 * aberr : invR5;
    const double crossSV = svx * ry - svy * rx;
    const double pMagMoment = p->magMoment, pAngMomentum = p->angMomentum;
    const int rad = s->rad;
    double vrx = 0, vry = 0, rDotVr = 0;                                  /* forces.js:198-203 */
    if (rad) { vrx = svx - p->vx; vry = svy - p->vy; rDotVr = rx * vrx + ry * vry; }

    if (s->grav) {                                                        /* forces.js:213-227 */
        const double k = p->mass * sMass;
        const double fDir = k * invR3a;
        p->fx += rx * fDir; p->fy += ry * fDir;
        if (rad) {
            const double jRadial = -3 * k * rDotVr * invR5a;
            p->jx += vrx * fDir + rx * jRadial; p->jy += vry * fDir + ry * jRadial;
        }
        if (accumulatePE) p->pe -= k * invR;
    }
    const double axModPair = s->needAxMod ? sqrt(p->axMod * sAxMod) : 1;  /* forces.js:230 */
    if (s->coul) {                                                        /* forces.js:232-246 */
        const double k = -(p->charge * sCharge) * axModPair;
        const double fDir = k * invR3a;
        p->fx += rx * fDir; p->fy += ry * fDir;
        if (rad) {
            const double jRadial = -3 * k * rDotVr * invR5a;
            p->jx += vrx * fDir + rx * jRadial; p->jy += vry * fDir + ry * jRadial;
        }
        if (accumulatePE) p->pe += p->charge * sCharge * invR * axModPair;
    }
    if (s->onePN && (s->gm || s->mag)) {                                  /* forces.js:249-301 */
        const double r = 1 / invR;
        const double nx = rx * invR, ny = ry * invR;
        const double pvx = p->vx, pvy = p->vy;
        const double v1DotV2 = pvx * svx + pvy * svy;
        if (s->gm) {
            const double v1Sq = pvx * pvx + pvy * pvy;
            const double v2Sq = svx * svx + svy * svy;
            const double nDotV1 = nx * pvx + ny * pvy;
            const double nDotV2 = nx * svx + ny * svy;
            const double radial = -v1Sq - 2 * v2Sq + 1.5 * nDotV2 * nDotV2
                                  + 5 * p->mass * invR + 4 * sMass * invR;
            const double v1Coeff = 4 * nDotV1 - 3 * nDotV2;
            const double v2Coeff = 3 * nDotV2;
            const double base = p->mass * sMass * invR3;
            const double fx = base * (rx * radial + (pvx * v1Coeff + svx * v2Coeff) * r);
            const double fy = base * (ry * radial + (pvy * v1Coeff + svy * v2Coeff) * r);
            p->fx += fx; p->fy += fy;
            p->f1x += fx; p->f1y += fy;
            if (rad) {                      /* position-only EIH term m1 m2 (5 m1 + 4 m2) r / r^4 */
                const double kEIH = p->mass * sMass * (5 * p->mass + 4 * sMass);
                const double fDirEIH = kEIH * invRSq * invRSq;
                const double jRadialEIH = -4 * kEIH * rDotVr * invRSq * invRSq * invRSq;
                p->jx += vrx * fDirEIH + rx * jRadialEIH; p->jy += vry * fDirEIH + ry * jRadialEIH;
            }
            if (accumulatePE)
                p->pe -= p->mass * sMass * invR * (1.5 * (v1Sq + v2Sq) - 3.5 * v1DotV2
                          - 0.5 * nDotV1 * nDotV2 + p->mass * invR + sMass * invR);
        }
        if (s->mag) {
            const double v2DotN = svx * nx + svy * ny;
            const double v1DotN = pvx * nx + pvy * ny;
            const double coeff = 0.5 * p->charge * sCharge * invRSq;
            const double symX = coeff * (pvx * v2DotN - 3 * nx * v1DotN * v2DotN);
            const double symY = coeff * (pvy * v2DotN - 3 * ny * v1DotN * v2DotN);
            p->fx += symX; p->fy += symY;
            p->f1x += symX; p->f1y += symY;
            if (accumulatePE) p->pe -= 0.5 * p->charge * sCharge * invR * (v1DotV2 + v1DotN * v2DotN);
        }
    }
    if (s->onePN && s->gm && s->mag) {                                    /* forces.js:303-320 */
        const double crossCoeff = p->charge * sCharge * (p->mass + sMass)
            - (p->charge * p->charge * sMass + sCharge * sCharge * p->mass);
        const double fDir = crossCoeff * invRSq * invRSq;
        p->fx += rx * fDir; p->fy += ry * fDir;
        p->f1x += rx * fDir; p->f1y += ry * fDir;
        if (rad) {
            const double jRadial = -4 * crossCoeff * rDotVr * invRSq * invRSq * invRSq;
            p->jx += vrx * fDir + rx * jRadial; p->jy += vry * fDir + ry * jRadial;
        }
        if (accumulatePE) p->pe += 0.5 * crossCoeff * invRSq;
    }
    if (s->mag) {                                                         /* forces.js:322-353 */
        const double axMod = axModPair;
        const double fDir = -3 * (pMagMoment * sMagMoment) * invR5a * axMod;
        p->fx += rx * fDir; p->fy += ry * fDir;
        if (rad) {
            const double invR7a = invR5a * invRSq;
            const double jRadial = 5 * 3 * (pMagMoment * sMagMoment) * rDotVr * invR7a * axMod;
            p->jx += vrx * fDir + rx * jRadial; p->jy += vry * fDir + ry * jRadial;
        }
        if (accumulatePE) p->pe += (pMagMoment * sMagMoment) * invR3 * axMod;
        const double BzMoving = sCharge * crossSV * invR3 * axMod;
        p->Bz += BzMoving;
        p->dBzdx += 3 * BzMoving * rx * invRSq + sCharge * svy * invR3 * axMod;
        p->dBzdy += 3 * BzMoving * ry * invRSq - sCharge * svx * invR3 * axMod;
        p->Bz -= sMagMoment * invR3 * axMod;
        p->dBzdx -= 3 * sMagMoment * rx * invR5 * axMod;
        p->dBzdy -= 3 * sMagMoment * ry * invR5 * axMod;
    }
    if (s->gm) {                                                          /* forces.js:355-387 */
        const double fDir = 3 * (pAngMomentum * sAngMomentum) * invR5a;
        p->fx += rx * fDir; p->fy += ry * fDir;
        if (rad) {
            const double invR7a = invR5a * invRSq;
            const double jRadial = -5 * 3 * (pAngMomentum * sAngMomentum) * rDotVr * invR7a;
            p->jx += vrx * fDir + rx * jRadial; p->jy += vry * fDir + ry * jRadial;
        }
        if (accumulatePE) p->pe -= (pAngMomentum * sAngMomentum) * invR3;
        const double BgzMoving = -sMass * crossSV * invR3;
        p->Bgz += BgzMoving;
        p->dBgzdx += 3 * BgzMoving * rx * invRSq - sMass * svy * invR3;
        p->dBgzdy += 3 * BgzMoving * ry * invRSq + sMass * svx * invR3;
        p->Bgz -= 2 * sAngMomentum * invR3;
        p->dBgzdx -= 6 * sAngMomentum * rx * invR5;
        p->dBgzdy -= 6 * sAngMomentum * ry * invR5;
        p->frameDrag += 2 * sAngMomentum * (sAngVel - p->angVel) * invR3;
    }
    if (s->yuk && rawRSq < s->yukCutoffSq) {                              /* forces.js:389-427 */
        const double mu = s->higgs ? g->yukawa_mu * sqrt(p->higgsMod * sHiggsMod) : g->yukawa_mu;
        const double r = 1 / invR;
        const double expMuR = exp(-mu * r);
        const double ym = s->needAxMod ? sqrt(p->yukMod * sYukMod) : 1;
        const double G = g->yukawa_coupling;
        const double fDir = G * ym * p->mass * sMass * expMuR * (invRSq + mu * invR) * (signalDelayed ? invR * aberr : invR);
        p->fx += rx * fDir; p->fy += ry * fDir;
        if (rad) {
            const double jRadial = -(3 * invRSq + 3 * mu * invR + mu * mu) * rDotVr * G * ym * p->mass * sMass * expMuR
                                   * invRSq * (signalDelayed ? invR * aberr : invR);
            p->jx += vrx * fDir + rx * jRadial; p->jy += vry * fDir + ry * jRadial;
        }
        if (accumulatePE) p->pe -= G * ym * p->mass * sMass * expMuR * invR;
        if (s->onePN) {
            const double pvx = p->vx, pvy = p->vy;
            const double nx = rx * invR, ny = ry * invR;
            const double nDotV1 = nx * pvx + ny * pvy;
            const double nDotV2 = nx * svx + ny * svy;
            const double v1DotV2 = pvx * svx + pvy * svy;
            const double alpha = 1 + mu * r;
            const double beta = 0.5 * G * ym * p->mass * sMass * expMuR * invRSq;
            const double radial = -(alpha * v1DotV2 + (alpha * alpha + alpha + 1) * nDotV1 * nDotV2);
            const double fx = beta * (radial * nx + alpha * (nDotV2 * pvx + nDotV1 * svx));
            const double fy = beta * (radial * ny + alpha * (nDotV2 * pvy + nDotV1 * svy));
            p->fx += fx; p->fy += fy;
            p->f1x += fx; p->f1y += fy;
            if (accumulatePE)
                p->pe += 0.5 * G * ym * p->mass * sMass * expMuR * invR * (v1DotV2 + nDotV1 * nDotV2 * alpha);
        }
    }
    if (s->grav) {                                                        /* forces.js:429-441 */
        const double crossRV = rx * (svy - p->vy) - ry * (svx - p->vx);
        const double wOrbit = crossRV * invRSq;
        const double dw = p->angVel - wOrbit;
        double coupling = sMass;
        if (s->coul && p->mass > EPS_) coupling += p->charge * sCharge / p->mass;
        const double bodyR = sqrt(p->bodyRadiusSq);
        const double ri5 = p->bodyRadiusSq * p->bodyRadiusSq * bodyR;
        const double invR6 = invRSq * invRSq * invRSq;
        p->tidal -= TIDAL_STRENGTH * coupling * coupling * ri5 * invR6 * dw;
    }
}

#include "geon_oracle_delay.inc"

/* ------------------------------------------------------------------ */
/* quadtree.js                                                         */
/* ------------------------------------------------------------------ */
static void pool_free(Pool* q) {
    free(q->bx); free(q->by); free(q->bw); free(q->bh);
    free(q->mass); free(q->charge); free(q->magMom); free(q->angMom);
    free(q->momX); free(q->momY); free(q->comX); free(q->comY); free(q->mom);
    free(q->nw); free(q->ne); free(q->sw); free(q->se);
    free(q->pointCount); free(q->divided); free(q->level); free(q->subCount); free(q->qq);
    free(q->path); free(q->points);
    memset(q, 0, sizeof(*q));
}

#define GROW(ptr, type, n) ptr = (type*)realloc(ptr, sizeof(type) * (size_t)(n))
static void pool_grow(Pool* q, int newCap) {                              /* quadtree.js:81-103 */
    GROW(q->bx, double, newCap); GROW(q->by, double, newCap);
    GROW(q->bw, double, newCap); GROW(q->bh, double, newCap);
    GROW(q->mass, double, newCap); GROW(q->charge, double, newCap);
    GROW(q->magMom, double, newCap); GROW(q->angMom, double, newCap);
    GROW(q->momX, double, newCap); GROW(q->momY, double, newCap);
    GROW(q->comX, double, newCap); GROW(q->comY, double, newCap);
    GROW(q->mom, double, (size_t)newCap * 9);
    GROW(q->nw, int32_t, newCap); GROW(q->ne, int32_t, newCap);
    GROW(q->sw, int32_t, newCap); GROW(q->se, int32_t, newCap);
    GROW(q->pointCount, int32_t, newCap); GROW(q->divided, int32_t, newCap);
    GROW(q->level, int32_t, newCap); GROW(q->subCount, int32_t, newCap);
    GROW(q->qq, int32_t, newCap); GROW(q->path, uint64_t, newCap);
    GROW(q->points, int32_t, (size_t)newCap * QT_CAP);
    q->cap = newCap;
}

static int pool_alloc(Pool* q, double bx, double by, double bw, double bh) {  /* quadtree.js:52-79 */
    if (q->count >= q->cap) pool_grow(q, q->cap ? q->cap * 2 : 512);
    int i = q->count++;
    q->bx[i] = bx; q->by[i] = by; q->bw[i] = bw; q->bh[i] = bh;
    q->mass[i] = q->charge[i] = q->magMom[i] = q->angMom[i] = 0;
    q->momX[i] = q->momY[i] = 0;
    q->comX[i] = bx; q->comY[i] = by;
    q->nw[i] = q->ne[i] = q->sw[i] = q->se[i] = -1;
    q->pointCount[i] = 0; q->divided[i] = 0;
    q->level[i] = 0; q->path[i] = 0; q->subCount[i] = 0; q->qq[i] = 0;
    return i;
}

static void pool_subdivide(Pool* q, int idx) {                           /* quadtree.js:119-172 */
    double x = q->bx[idx], y = q->by[idx];
    double hw = q->bw[idx] / 2, hh = q->bh[idx] / 2;
    if (q->count + 4 > q->cap) pool_grow(q, q->cap * 2);
    int c = q->count;
    q->count = c + 4;
    const double cx[4] = {x - hw, x + hw, x - hw, x + hw};
    const double cy[4] = {y - hh, y - hh, y + hh, y + hh};
    for (int k = 0; k < 4; k++) {
        int j = c + k;
        q->bx[j] = cx[k]; q->by[j] = cy[k]; q->bw[j] = hw; q->bh[j] = hh;
        q->mass[j] = q->charge[j] = q->magMom[j] = q->angMom[j] = 0;
        q->momX[j] = q->momY[j] = 0;
        q->comX[j] = cx[k]; q->comY[j] = cy[k];
        q->nw[j] = q->ne[j] = q->sw[j] = q->se[j] = -1;
        q->pointCount[j] = 0; q->divided[j] = 0;
        q->level[j] = q->level[idx] + 1;
        q->path[j] = (q->path[idx] << 2) | (uint64_t)k;
        q->subCount[j] = 0; q->qq[j] = 0;
    }
    q->nw[idx] = c; q->ne[idx] = c + 1; q->sw[idx] = c + 2; q->se[idx] = c + 3;
    q->divided[idx] = 1;
}

static inline int pool_childFor(const Pool* q, int idx, double px, double py) {  /* quadtree.js:175-179 */
    return py <= q->by[idx] ? (px <= q->bx[idx] ? q->nw[idx] : q->ne[idx])
                            : (px <= q->bx[idx] ? q->sw[idx] : q->se[idx]);
}

static inline int pool_contains(const Pool* q, int idx, double px, double py) {  /* quadtree.js:105-110 */
    return px >= q->bx[idx] - q->bw[idx] && px <= q->bx[idx] + q->bw[idx] &&
           py >= q->by[idx] - q->bh[idx] && py <= q->by[idx] + q->bh[idx];
}

/* Iterative insert with a (node, point) work stack (quadtree.js:185-237). */
static int pool_insert(Pool* q, int idx, const P* tp, int pi, int32_t** stkN, int32_t** stkP, int* stkCap) {
    if (!pool_contains(q, idx, tp[pi].x, tp[pi].y)) return 0;
    int top = 0;
    if (*stkCap < 64) { *stkCap = 64; *stkN = realloc(*stkN, 64 * 4); *stkP = realloc(*stkP, 64 * 4); }
    (*stkN)[top] = idx; (*stkP)[top] = pi; top++;
    while (top > 0) {
        top--;
        int nodeIdx = (*stkN)[top];
        int pt = (*stkP)[top];
        const double px = tp[pt].x, py = tp[pt].y;
        int depth = 0;
        while (depth <= 48) {
            if (q->pointCount[nodeIdx] < QT_CAP && !q->divided[nodeIdx]) {
                q->points[(size_t)nodeIdx * QT_CAP + q->pointCount[nodeIdx]] = pt;
                q->pointCount[nodeIdx]++;
                break;
            }
            if (!q->divided[nodeIdx]) {
                pool_subdivide(q, nodeIdx);
                if (top + QT_CAP + 1 > *stkCap) {
                    *stkCap *= 2;
                    *stkN = realloc(*stkN, (size_t)*stkCap * 4);
                    *stkP = realloc(*stkP, (size_t)*stkCap * 4);
                }
                size_t base = (size_t)nodeIdx * QT_CAP;
                for (int i = 0; i < q->pointCount[nodeIdx]; i++) {
                    int pp = q->points[base + i];
                    (*stkN)[top] = pool_childFor(q, nodeIdx, tp[pp].x, tp[pp].y);
                    (*stkP)[top] = pp;
                    top++;
                }
                q->pointCount[nodeIdx] = 0;
            }
            nodeIdx = pool_childFor(q, nodeIdx, px, py);
            depth++;
        }
    }
    return 1;
}

/* Post-order FP64 aggregates (quadtree.js:239-312); also subtree point counts
 * and exact charge-quanta sums for the canonical export. */
static void pool_massDistribution(Pool* q, const P* tp, int root) {
    int* order = (int*)malloc(sizeof(int) * (size_t)q->count);
    int* stack = (int*)malloc(sizeof(int) * (size_t)(q->count + 4));
    int top = 0, len = 0;
    stack[top++] = root;
    while (top > 0) {
        int idx = stack[--top];
        order[len++] = idx;
        if (q->divided[idx]) {
            stack[top++] = q->nw[idx]; stack[top++] = q->ne[idx];
            stack[top++] = q->sw[idx]; stack[top++] = q->se[idx];
        }
    }
    for (int k = len - 1; k >= 0; k--) {
        int idx = order[k];
        if (!q->divided[idx]) {
            int cnt = q->pointCount[idx];
            q->subCount[idx] = cnt;
            if (cnt == 0) continue;
            double mass = 0, charge = 0, magMom = 0, angMom = 0, cx = 0, cy = 0, momX = 0, momY = 0;
            int qq = 0;
            size_t base = (size_t)idx * QT_CAP;
            for (int i = 0; i < cnt; i++) {
                const P* p = &tp[q->points[base + i]];
                double pm = p->mass;
                mass += pm; charge += p->charge; magMom += p->magMoment; angMom += p->angMomentum;
                cx += p->x * pm; cy += p->y * pm;
                momX += pm * p->wx; momY += pm * p->wy;
                qq += p->q_quanta;
            }
            q->mass[idx] = mass; q->charge[idx] = charge;
            q->magMom[idx] = magMom; q->angMom[idx] = angMom;
            q->momX[idx] = momX; q->momY[idx] = momY;
            q->qq[idx] = qq;
            if (mass > 0) { q->comX[idx] = cx / mass; q->comY[idx] = cy / mass; }
        } else {
            int c0 = q->nw[idx], c1 = q->ne[idx], c2 = q->sw[idx], c3 = q->se[idx];
            double m0 = q->mass[c0], m1 = q->mass[c1], m2 = q->mass[c2], m3 = q->mass[c3];
            double mass = m0 + m1 + m2 + m3;
            q->mass[idx] = mass;
            q->charge[idx] = q->charge[c0] + q->charge[c1] + q->charge[c2] + q->charge[c3];
            q->magMom[idx] = q->magMom[c0] + q->magMom[c1] + q->magMom[c2] + q->magMom[c3];
            q->angMom[idx] = q->angMom[c0] + q->angMom[c1] + q->angMom[c2] + q->angMom[c3];
            q->momX[idx] = q->momX[c0] + q->momX[c1] + q->momX[c2] + q->momX[c3];
            q->momY[idx] = q->momY[c0] + q->momY[c1] + q->momY[c2] + q->momY[c3];
            q->subCount[idx] = q->subCount[c0] + q->subCount[c1] + q->subCount[c2] + q->subCount[c3];
            q->qq[idx] = q->qq[c0] + q->qq[c1] + q->qq[c2] + q->qq[c3];
            if (mass > 0) {
                q->comX[idx] = (q->comX[c0] * m0 + q->comX[c1] * m1 + q->comX[c2] * m2 + q->comX[c3] * m3) / mass;
                q->comY[idx] = (q->comY[c0] * m0 + q->comY[c1] * m1 + q->comY[c2] * m2 + q->comY[c3] * m3) / mass;
            }
        }
    }
    free(order); free(stack);
}

/* Second moments about each node's COM (not in the reference: the input of the
 * far-field correction).  Leaves sum their points; a divided node combines its
 * children by the parallel-axis shifts:
 *   Qm += Qm_c + m_c D D^T,   Dq += Dq_c + q_c D,
 *   Qq += Qq_c + Dq_c D^T + D Dq_c^T + q_c D D^T   (D = com_c - com). */
static void pool_moments(Pool* q, const P* tp, int root) {
    int* order = (int*)malloc(sizeof(int) * (size_t)q->count);
    int* stack = (int*)malloc(sizeof(int) * (size_t)(q->count + 4));
    int top = 0, len = 0;
    stack[top++] = root;
    while (top > 0) {
        int idx = stack[--top];
        order[len++] = idx;
        if (q->divided[idx]) {
            stack[top++] = q->nw[idx]; stack[top++] = q->ne[idx];
            stack[top++] = q->sw[idx]; stack[top++] = q->se[idx];
        }
    }
    for (int k = len - 1; k >= 0; k--) {
        const int idx = order[k];
        double* M = &q->mom[(size_t)idx * 9];
        for (int c = 0; c < 9; c++) M[c] = 0;
        const double cx = q->comX[idx], cy = q->comY[idx];
        if (!q->divided[idx]) {
            const size_t base = (size_t)idx * QT_CAP;
            for (int i = 0; i < q->pointCount[idx]; i++) {
                const P* p = &tp[q->points[base + i]];
                const double dx = p->x - cx, dy = p->y - cy;
                M[0] += p->mass * dx * dx; M[1] += p->mass * dx * dy; M[2] += p->mass * dy * dy;
                M[3] += p->charge * dx; M[4] += p->charge * dy;
                M[5] += p->charge * dx * dx; M[6] += p->charge * dx * dy; M[7] += p->charge * dy * dy;
                M[8] += fabs(p->charge);
            }
        } else {
            const int ch[4] = {q->nw[idx], q->ne[idx], q->sw[idx], q->se[idx]};
            for (int c = 0; c < 4; c++) {
                const int j = ch[c];
                const double* C = &q->mom[(size_t)j * 9];
                const double m = q->mass[j], qc = q->charge[j];
                const double dx = q->comX[j] - cx, dy = q->comY[j] - cy;
                if (m <= 0 && C[8] == 0) continue;
                M[0] += C[0] + m * dx * dx; M[1] += C[1] + m * dx * dy; M[2] += C[2] + m * dy * dy;
                M[3] += C[3] + qc * dx; M[4] += C[4] + qc * dy;
                M[5] += C[5] + 2 * C[3] * dx + qc * dx * dx;
                M[6] += C[6] + C[3] * dy + C[4] * dx + qc * dx * dy;
                M[7] += C[7] + 2 * C[4] * dy + qc * dy * dy;
                M[8] += C[8];
            }
        }
    }
    free(order); free(stack);
}

/* Far-field correction of an accepted node (not in the reference): the
 * gravity and Coulomb force and potential terms beyond the monopole of the softened kernel
 * g(d) = d (|d|^2 + eps^2)^(-3/2), with coefficients c[0..7] = mass quadrupole,
 * charge dipole, charge quadrupole (about the COM):
 *   1/2 Q:dd g = 2 (Q d) h' + tr(Q) d h' + 2 d (d.Q.d) h'',  D.dg = D h + 2 d (D.d) h',
 * h = r2^-3/2, h' = -3/2 r2^-5/2, h'' = 15/4 r2^-7/2, r2 = |d|^2 + eps^2. */
static void far_correction(const OrcSys* s, P* p, double dx, double dy, const double* c) {
    const double r2 = dx * dx + dy * dy + s->softeningSq;
    const double ir2 = 1 / r2, ir = sqrt(ir2);
    const double h = ir * ir2, h1 = -1.5 * h * ir2, h2 = 3.75 * h * ir2 * ir2;
    double fx = 0, fy = 0;
    if (s->grav) {
        const double qx = c[0] * dx + c[1] * dy, qy = c[1] * dx + c[2] * dy;
        const double dqd = dx * qx + dy * qy, tr = c[0] + c[2];
        fx += p->mass * (2 * qx * h1 + tr * dx * h1 + 2 * dx * dqd * h2);
        fy += p->mass * (2 * qy * h1 + tr * dy * h1 + 2 * dy * dqd * h2);
        p->pe += p->mass * (0.5 * tr * h + dqd * h1);           /* -m_i (1/2) Q : grad grad r2^-1/2 */
    }
    if (s->coul) {
        const double dd = c[3] * dx + c[4] * dy;
        const double qx = c[5] * dx + c[6] * dy, qy = c[6] * dx + c[7] * dy;
        const double dqd = dx * qx + dy * qy, tr = c[5] + c[7];
        fx -= p->charge * (c[3] * h + 2 * dx * dd * h1 + 2 * qx * h1 + tr * dx * h1 + 2 * dx * dqd * h2);
        fy -= p->charge * (c[4] * h + 2 * dy * dd * h1 + 2 * qy * h1 + tr * dy * h1 + 2 * dy * dqd * h2);
        p->pe -= p->charge * (dd * h + 0.5 * tr * h + dqd * h1);
    }
    p->fx += fx; p->fy += fy;
}

/* ------------------------------------------------------------------ */
/* Learned far-field correction: FP64 restatement of the network of      */
/* physsim_amd/csrc/k_learn.hip (not in the reference).                  */
/* ------------------------------------------------------------------ */
#define FAR_IN 10
#define FAR_HID 64
#define FAR_W1 0
#define FAR_W2 1024
#define FAR_B2 5120
#define FAR_OUT 6
#define FAR_W3 5184
#define FAR_B3 5568
#define FAR_PARAMS 5574
#define FAR_MODEL 5600
#define FAR_SAMPLE 20
#define FAR_WSEL(j) ((j) < 4 ? (j) >> 1 : (j) - 2)

/* Hard-swish activation of the network: z clamp((z + 3)/6, 0, 1). */
static double act64(double z) {
    const double t = (z + 3.0) / 6.0;
    return z * (t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t));
}

static double act64_grad(double z) { return z < -3.0 ? 0.0 : (z > 3.0 ? 1.0 : (2.0 * z + 3.0) / 6.0); }

/* Features of an (observer, node) pair in the offset frame u = d/|d|,
 * v = (-u_y, u_x); c = the oracle's 9 node moments (Qm xx xy yy, Dq x y,
 * Qq xx xy yy, sum |q|).  sg0/sc0: the output scales without m_i, q_i. */
static void far_features(double dx, double dy, double eps2, double s, double M, const double* c, double* f,
                         double* ux, double* uy, double* sg0, double* sc0, double* sp0, double* cp0) {
    const double rho2 = dx * dx + dy * dy, r2 = rho2 + eps2, r = sqrt(r2), irho = 1 / sqrt(rho2);
    const double x = dx * irho, y = dy * irho;
    const double sr = s / r;
    const double iM = M > 0 ? 1 / (M * s * s) : 0, A = c[8], iA = A > 0 ? 1 / (A * s) : 0;
    const double uu = x * x, vv = y * y, uv = x * y;
    f[0] = sr;
    f[1] = eps2 / r2;
    f[2] = (uu * c[0] + 2 * uv * c[1] + vv * c[2]) * iM;
    f[3] = (-uv * c[0] + (uu - vv) * c[1] + uv * c[2]) * iM;
    f[4] = (vv * c[0] - 2 * uv * c[1] + uu * c[2]) * iM;
    f[5] = (x * c[3] + y * c[4]) * iA;
    f[6] = (-y * c[3] + x * c[4]) * iA;
    f[7] = (uu * c[5] + 2 * uv * c[6] + vv * c[7]) * iA / s;
    f[8] = (-uv * c[5] + (uu - vv) * c[6] + uv * c[7]) * iA / s;
    f[9] = (vv * c[5] - 2 * uv * c[6] + uu * c[7]) * iA / s;
    *ux = x; *uy = y;
    *sg0 = M * sr * sr / r2;
    *sc0 = A * sr / r2;
    *sp0 = M * sr * sr / r;
    *cp0 = A * sr / r;
}

/* Forward pass; keeps the pre-activations when z1/z2 are given. */
static void far_forward(const double* P, const double* feat, double* x, double* z1, double* a1, double* z2,
                        double* a2, double* o) {
    const double* mean = P + FAR_PARAMS;
    const double* istd = mean + FAR_IN;
    for (int k = 0; k < 16; k++) x[k] = 0;
    for (int k = 0; k < FAR_IN; k++) x[k] = (feat[k] - mean[k]) * istd[k];
    x[FAR_IN] = 1;
    for (int m = 0; m < FAR_HID; m++) {
        double z = 0;
        for (int k = 0; k < 16; k++) z += x[k] * P[FAR_W1 + k * FAR_HID + m];
        z1[m] = z; a1[m] = act64(z);
    }
    for (int b = 0; b < FAR_HID; b++) {
        double z = P[FAR_B2 + b];
        for (int a = 0; a < FAR_HID; a++) z += a1[a] * P[FAR_W2 + a * FAR_HID + b];
        z2[b] = z; a2[b] = act64(z);
    }
    for (int j = 0; j < FAR_OUT; j++) {
        double z = P[FAR_B3 + j];
        for (int b = 0; b < FAR_HID; b++) z += a2[b] * P[FAR_W3 + b * FAR_OUT + j];
        o[j] = z;
    }
}

/* The learned residual of an accepted node, added to the observer's force. */
static void far_learned(const OrcSys* s, P* p, double dx, double dy, int ni) {
    const Pool* q = &s->pool;
    const double* P = s->farModel;
    double f[FAR_IN], ux, uy, sg0, sc0, sp0, cp0, x[16], z1[FAR_HID], a1[FAR_HID], z2[FAR_HID], a2[FAR_HID], o[FAR_OUT];
    far_features(dx, dy, s->softeningSq, q->bw[ni] * 2, q->mass[ni], &q->mom[(size_t)ni * 9], f, &ux, &uy, &sg0, &sc0,
                 &sp0, &cp0);
    far_forward(P, f, x, z1, a1, z2, a2, o);
    const double* os = P + FAR_PARAMS + 2 * FAR_IN;
    const double gu = o[0] * os[0], gv = o[1] * os[1], cu = o[2] * os[2], cv = o[3] * os[3];
    const double sg = s->grav ? p->mass * sg0 : 0, sc = s->coul ? p->charge * sc0 : 0;
    const double sp = s->grav ? p->mass * sp0 : 0, cp = s->coul ? p->charge * cp0 : 0;
    p->fx += sg * (ux * gu - uy * gv) + sc * (ux * cu - uy * cv);
    p->fy += sg * (uy * gu + ux * gv) + sc * (uy * cu + ux * cv);
    p->pe += sp * o[4] * os[4] + cp * o[5] * os[5];
}

int orc_far_set_model(OrcSys* s, const float* model) {
    if (!s->farModel) s->farModel = (double*)malloc(sizeof(double) * FAR_MODEL);
    for (int i = 0; i < FAR_MODEL; i++) s->farModel[i] = model[i];
    return 0;
}

/* Weighted squared-error loss of the network on n samples and its gradient
 * (FAR_PARAMS doubles), in FP64:
 *   L = wscale / n * sum_r sum_j w_rj (oscale_j o_rj - t_rj)^2. */
double orc_far_loss_grad(const float* model, const float* samples, int64_t n, double wscale, double* grad) {
    double P[FAR_MODEL];
    for (int i = 0; i < FAR_MODEL; i++) P[i] = model[i];
    const double* os = P + FAR_PARAMS + 2 * FAR_IN;
    for (int i = 0; i < FAR_PARAMS; i++) grad[i] = 0;
    double L = 0;
    for (int64_t r = 0; r < n; r++) {
        const float* sm = samples + r * FAR_SAMPLE;
        double feat[FAR_IN], x[16], z1[FAR_HID], a1[FAR_HID], z2[FAR_HID], a2[FAR_HID], o[FAR_OUT], dO[FAR_OUT];
        for (int k = 0; k < FAR_IN; k++) feat[k] = sm[k];
        far_forward(P, feat, x, z1, a1, z2, a2, o);
        for (int j = 0; j < FAR_OUT; j++) {
            const double w = wscale * sm[FAR_IN + FAR_OUT + FAR_WSEL(j)];
            const double err = os[j] * o[j] - sm[FAR_IN + j];
            L += w * err * err / n;
            dO[j] = 2 * w * err * os[j] / n;
        }
        double dz2[FAR_HID], dz1[FAR_HID];
        for (int b = 0; b < FAR_HID; b++) {
            double da = 0;
            for (int j = 0; j < FAR_OUT; j++) {
                da += P[FAR_W3 + b * FAR_OUT + j] * dO[j];
                grad[FAR_W3 + b * FAR_OUT + j] += a2[b] * dO[j];
            }
            dz2[b] = da * act64_grad(z2[b]);
            grad[FAR_B2 + b] += dz2[b];
        }
        for (int j = 0; j < FAR_OUT; j++) grad[FAR_B3 + j] += dO[j];
        for (int a = 0; a < FAR_HID; a++) {
            double da = 0;
            for (int b = 0; b < FAR_HID; b++) { da += P[FAR_W2 + a * FAR_HID + b] * dz2[b]; grad[FAR_W2 + a * FAR_HID + b] += a1[a] * dz2[b]; }
            dz1[a] = da * act64_grad(z1[a]);
        }
        for (int k = 0; k < 16; k++)
            for (int m = 0; m < FAR_HID; m++) grad[FAR_W1 + k * FAR_HID + m] += x[k] * dz1[m];
    }
    return L;
}

/* ------------------------------------------------------------------ */
/* integrator.js: ghosts and tree build                                */
/* ------------------------------------------------------------------ */
static void tp_reserve(OrcSys* s, int n) {
    if (n > s->tpcap) {
        s->tpcap = n * 2 + 16;
        s->tp = (P*)realloc(s->tp, sizeof(P) * (size_t)s->tpcap);
    }
}

static void addGhost(OrcSys* s, int pi, double sx, double sy, int flipVx, int flipVy) {  /* integrator.js:178-198 */
    const P* p = &s->p[pi];
    P* g = &s->tp[s->ntp++];
    *g = *p;
    g->x = sx; g->y = sy;
    g->vx = flipVx ? -p->vx : p->vx; g->vy = flipVy ? -p->vy : p->vy;
    g->wx = flipVx ? -p->wx : p->wx; g->wy = flipVy ? -p->wy : p->wy;
    g->angVel = (flipVx || flipVy) ? -p->angVel : p->angVel;
    g->angw = (flipVx || flipVy) ? -p->angw : p->angw;
    g->magMoment = MAG_MOMENT_K * g->charge * g->angVel * g->radiusSq;
    g->angMomentum = INERTIA_K * g->mass * g->angVel * g->radiusSq;
    g->isGhost = 1; g->original = pi; g->id = -1;
}

static void generateGhosts(OrcSys* s) {                                    /* integrator.js:201-244 */
    const double W = s->prm.domain_w, H = s->prm.domain_h;
    const double margin = (W > H ? W : H) * s->prm.theta;
    const int topo = s->prm.topology;
    for (int i = 0; i < s->n; i++) {
        const double x = s->p[i].x, y = s->p[i].y;
        int nearL = x < margin, nearR = x > W - margin;
        int nearT = y < margin, nearB = y > H - margin;
        tp_reserve(s, s->ntp + 8);
        if (topo == GEON_TORUS) {
            if (nearL) addGhost(s, i, x + W, y, 0, 0);
            if (nearR) addGhost(s, i, x - W, y, 0, 0);
            if (nearT) addGhost(s, i, x, y + H, 0, 0);
            if (nearB) addGhost(s, i, x, y - H, 0, 0);
            if (nearL && nearT) addGhost(s, i, x + W, y + H, 0, 0);
            if (nearL && nearB) addGhost(s, i, x + W, y - H, 0, 0);
            if (nearR && nearT) addGhost(s, i, x - W, y + H, 0, 0);
            if (nearR && nearB) addGhost(s, i, x - W, y - H, 0, 0);
        } else if (topo == GEON_KLEIN) {
            if (nearL) addGhost(s, i, x + W, y, 0, 0);
            if (nearR) addGhost(s, i, x - W, y, 0, 0);
            if (nearT) addGhost(s, i, W - x, y + H, 1, 0);
            if (nearB) addGhost(s, i, W - x, y - H, 1, 0);
            if (nearL && nearT) addGhost(s, i, W - x + W, y + H, 1, 0);
            if (nearL && nearB) addGhost(s, i, W - x + W, y - H, 1, 0);
            if (nearR && nearT) addGhost(s, i, W - x - W, y + H, 1, 0);
            if (nearR && nearB) addGhost(s, i, W - x - W, y - H, 1, 0);
        } else {
            if (nearL) addGhost(s, i, x + W, H - y, 0, 1);
            if (nearR) addGhost(s, i, x - W, H - y, 0, 1);
            if (nearT) addGhost(s, i, W - x, y + H, 1, 0);
            if (nearB) addGhost(s, i, W - x, y - H, 1, 0);
            if (nearL && nearT) addGhost(s, i, W - x + W, y + H, 1, 0);
            if (nearL && nearB) addGhost(s, i, W - x + W, y - H, 1, 0);
            if (nearR && nearT) addGhost(s, i, W - x - W, y + H, 1, 0);
            if (nearR && nearB) addGhost(s, i, W - x - W, y - H, 1, 0);
        }
    }
}

/* _buildTree (integrator.js:247-260) + QuadTreePool.build (quadtree.js:412-419).
 * Root box: centre (W/2, H/2), half-extent (2W, 2H) (integrator.js:727-730). */
/* ghosts: the periodic images of _buildTree (integrator.js:247-258) go into
 * the tree; the force walk takes them only in GEON_PERIODIC_GHOSTS mode. */
static int buildTreeG(OrcSys* s, int ghosts) {
    tp_reserve(s, s->n + 8);
    s->ntp = 0;
    for (int i = 0; i < s->n; i++) { s->tp[s->ntp] = s->p[i]; s->tp[s->ntp].isGhost = 0; s->tp[s->ntp].original = i; s->ntp++; }
    if (ghosts) generateGhosts(s);
    Pool* q = &s->pool;
    q->count = 0;
    const double W = s->prm.domain_w, H = s->prm.domain_h;
    int root = pool_alloc(q, W / 2, H / 2, W * 2, H * 2);
    int32_t *stkN = NULL, *stkP = NULL; int stkCap = 0;
    for (int i = 0; i < s->ntp; i++) pool_insert(q, root, s->tp, i, &stkN, &stkP, &stkCap);
    free(stkN); free(stkP);
    pool_massDistribution(q, s->tp, root);
    pool_moments(q, s->tp, root);
    s->root = root;
    return root;
}

static int buildTree(OrcSys* s) { return buildTreeG(s, s->periodic && s->prm.periodic_mode == GEON_PERIODIC_GHOSTS); }

/* ------------------------------------------------------------------ */
/* forces.js: walk, direct sum, computeAllForces                       */
/* ------------------------------------------------------------------ */
/* GEON_PERIODIC_WINDOWED: a node may stand for its particles only if its whole
 * box, moved by the image shift of its COM, lies inside the observer's
 * min-image window (not in the reference: see include/geon.h). */
static inline int window_ok(const OrcSys* s, const Pool* q, int ni, double px, double py, double dx, double dy) {
    if (!s->periodic || s->prm.periodic_mode != GEON_PERIODIC_WINDOWED) return 1;
    const double W = s->prm.domain_w, H = s->prm.domain_h;
    const double sx = dx - (q->comX[ni] - px), sy = dy - (q->comY[ni] - py);
    const double cx = q->bx[ni] - px + sx, cy = q->by[ni] - py + sy;
    return fabs(cx) + q->bw[ni] <= 0.5 * W && fabs(cy) + q->bh[ni] <= 0.5 * H;
}

/* calculateForce (forces.js:622-694); leaf sources from their past light cone
 * when signal delay is on (651-672), accepted nodes at the current time. */
static void calculateForce(OrcSys* s, int pi, int32_t* stack, int64_t* cnt) {
    const Pool* q = &s->pool;
    P* p = &s->p[pi];
    const double thetaSq = s->prm.theta * s->prm.theta;
    const double W = s->prm.domain_w, H = s->prm.domain_h;
    const double px = p->x, py = p->y;
    int top = 0;
    stack[top++] = s->root;
    while (top > 0) {
        int nodeIdx = stack[--top];
        cnt[2]++;
        if (q->mass[nodeIdx] < EPS_) continue;
        double dx, dy;
        if (s->periodic) minImage(px, py, q->comX[nodeIdx], q->comY[nodeIdx], s->prm.topology, W, H, W * 0.5, H * 0.5, &dx, &dy);
        else { dx = q->comX[nodeIdx] - px; dy = q->comY[nodeIdx] - py; }
        const double dSq = dx * dx + dy * dy;
        const double size = q->bw[nodeIdx] * 2;
        if (!q->divided[nodeIdx] && q->pointCount[nodeIdx] > 0) {
            size_t base = (size_t)nodeIdx * QT_CAP;
            for (int i = 0; i < q->pointCount[nodeIdx]; i++) {
                int oi = q->points[base + i];
                const P* o = &s->tp[oi];
                if (o->original == pi) continue;    /* other === particle, or its own ghost */
                const P* real = &s->p[o->original];
                if (!o->isGhost) o = real;          /* a real source is the live object (fresh μ, L) */
                if (s->delay) {
                    Delayed d;
                    if (!delayedState(s, real, px, py, &d)) continue;
                    const double sx = d.x + (o->x - real->x), sy = d.y + (o->y - real->y);   /* ghost shift */
                    const double r2 = real->bodyRadiusSq;
                    const double av = d.angw / sqrt(1 + d.angw * d.angw * r2);
                    pairForce(s, p, sx, sy, d.vx, d.vy, o->mass, o->charge, av, MAG_MOMENT_K * o->charge * av * r2,
                              INERTIA_K * o->mass * av * r2, real->axMod, real->yukMod, real->higgsMod, 1, 1);
                } else {
                    pairForce(s, p, o->x, o->y, o->vx, o->vy, o->mass, o->charge, o->angVel,
                              o->magMoment, o->angMomentum, real->axMod, real->yukMod, real->higgsMod, 1, 0);
                }
                cnt[0]++;
            }
        } else if (q->divided[nodeIdx] && (size * size < thetaSq * dSq) && window_ok(s, q, nodeIdx, px, py, dx, dy)) {
            const double nodeMass = q->mass[nodeIdx];
            pairForce(s, p, q->comX[nodeIdx], q->comY[nodeIdx], q->momX[nodeIdx] / nodeMass,
                      q->momY[nodeIdx] / nodeMass, nodeMass, q->charge[nodeIdx], 0,
                      q->magMom[nodeIdx], q->angMom[nodeIdx], 1, 1, 1, 1, 0);
            if (s->prm.far_field == GEON_FAR_MULTIPOLE) far_correction(s, p, dx, dy, &q->mom[(size_t)nodeIdx * 9]);
            if (s->prm.far_field == GEON_FAR_LEARNED && s->farModel) far_learned(s, p, dx, dy, nodeIdx);
            cnt[1]++;
        } else if (q->divided[nodeIdx]) {
            stack[top++] = q->nw[nodeIdx]; stack[top++] = q->ne[nodeIdx];
            stack[top++] = q->sw[nodeIdx]; stack[top++] = q->se[nodeIdx];
        }
    }
}

static void syncToggles(OrcSys* s);

/* Sum over the points under node ni of the softened gravity (per unit observer
 * mass) and Coulomb (per unit observer charge) kernels, each point placed at
 * offset (dx, dy) + (x_j - com) from the observer: the exact counterpart of
 * the node's monopole. */
static void node_exact(const OrcSys* s, const Pool* q, int ni, double dx, double dy, double cx, double cy,
                       int32_t* stack, double* g) {
    int top = 0;
    stack[top++] = ni;
    g[0] = g[1] = g[2] = g[3] = g[4] = g[5] = 0;
    while (top > 0) {
        const int k = stack[--top];
        if (q->divided[k]) {
            stack[top++] = q->nw[k]; stack[top++] = q->ne[k];
            stack[top++] = q->sw[k]; stack[top++] = q->se[k];
            continue;
        }
        const size_t base = (size_t)k * QT_CAP;
        for (int i = 0; i < q->pointCount[k]; i++) {
            const P* o = &s->tp[q->points[base + i]];
            const double ex = dx + (o->x - cx), ey = dy + (o->y - cy);
            const double r2 = ex * ex + ey * ey + s->softeningSq;
            const double ir = 1 / sqrt(r2), h = ir / r2;
            g[0] += o->mass * ex * h; g[1] += o->mass * ey * h;
            g[2] -= o->charge * ex * h; g[3] -= o->charge * ey * h;
            g[4] -= o->mass * ir; g[5] += o->charge * ir;
        }
    }
}

int64_t orc_far_samples(OrcSys* s, const int32_t* idx, int32_t n_idx, double* out, int64_t cap) {
    syncToggles(s);
    buildTree(s);
    const Pool* q = &s->pool;
    const double thetaSq = s->prm.theta * s->prm.theta;
    const double W = s->prm.domain_w, H = s->prm.domain_h;
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)(q->count * 2 + 16));
    int32_t* stack2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(q->count * 2 + 16));
    int64_t m = 0;
    for (int t = 0; t < n_idx; t++) {
        const int pi = idx[t];
        const double px = s->p[pi].x, py = s->p[pi].y;
        int top = 0;
        stack[top++] = s->root;
        while (top > 0) {                   /* the acceptance of calculateForce */
            const int ni = stack[--top];
            if (q->mass[ni] < EPS_ || !q->divided[ni]) continue;
            double dx, dy;
            if (s->periodic) minImage(px, py, q->comX[ni], q->comY[ni], s->prm.topology, W, H, W * 0.5, H * 0.5, &dx, &dy);
            else { dx = q->comX[ni] - px; dy = q->comY[ni] - py; }
            const double size = q->bw[ni] * 2;
            if (size * size < thetaSq * (dx * dx + dy * dy) && window_ok(s, q, ni, px, py, dx, dy)) {
                if (out && m < cap) {
                    double* r = out + (size_t)m * ORC_FAR_STRIDE;
                    double g[6];
                    node_exact(s, q, ni, dx, dy, q->comX[ni], q->comY[ni], stack2, g);
                    const double r2 = dx * dx + dy * dy + s->softeningSq;
                    const double ir = 1 / sqrt(r2), h = ir / r2;
                    r[0] = pi; r[1] = dx; r[2] = dy; r[3] = r2; r[4] = size;
                    r[5] = q->mass[ni]; r[6] = q->charge[ni];
                    for (int c = 0; c < 9; c++) r[7 + c] = q->mom[(size_t)ni * 9 + c];
                    r[16] = g[0] - q->mass[ni] * dx * h; r[17] = g[1] - q->mass[ni] * dy * h;
                    r[18] = g[2] + q->charge[ni] * dx * h; r[19] = g[3] + q->charge[ni] * dy * h;
                    r[20] = g[4] + q->mass[ni] * ir; r[21] = g[5] - q->charge[ni] * ir;
                }
                m++;
            } else {
                stack[top++] = q->nw[ni]; stack[top++] = q->ne[ni];
                stack[top++] = q->sw[ni]; stack[top++] = q->se[ni];
            }
        }
    }
    free(stack); free(stack2);
    return m;
}

/* Training samples in the GEON_FAR_SAMPLE layout (features, targets in the
 * offset frame normalised by the output scales, loss weights divided by f2ref)
 * for the observers in idx; obs[k] = the observer of sample k. */
int64_t orc_far_training_samples(OrcSys* s, const int32_t* idx, int32_t n_idx, double f2ref, double pe2ref,
                                 float* out, int32_t* obs, int64_t cap) {
    const int64_t m = orc_far_samples(s, idx, n_idx, NULL, 0);
    double* raw = (double*)malloc(sizeof(double) * ORC_FAR_STRIDE * (size_t)(m > 0 ? m : 1));
    orc_far_samples(s, idx, n_idx, raw, m);
    for (int64_t k = 0; k < m && k < cap; k++) {
        const double* r = raw + (size_t)k * ORC_FAR_STRIDE;
        const P* p = &s->p[(int)r[0]];
        double f[FAR_IN], ux, uy, sg0, sc0, sp0, cp0;
        far_features(r[1], r[2], s->softeningSq, r[4], r[5], r + 7, f, &ux, &uy, &sg0, &sc0, &sp0, &cp0);
        float* o = out + (size_t)k * FAR_SAMPLE;
        for (int j = 0; j < FAR_IN; j++) o[j] = (float)f[j];
        const double ig = sg0 > 0 ? 1 / sg0 : 0, ic = sc0 > 0 ? 1 / sc0 : 0;
        const double igp = sp0 > 0 ? 1 / sp0 : 0, icp = cp0 > 0 ? 1 / cp0 : 0;
        float* t = o + FAR_IN;
        t[0] = (float)((ux * r[16] + uy * r[17]) * ig);
        t[1] = (float)((-uy * r[16] + ux * r[17]) * ig);
        t[2] = (float)((ux * r[18] + uy * r[19]) * ic);
        t[3] = (float)((-uy * r[18] + ux * r[19]) * ic);
        t[4] = (float)(r[20] * igp);
        t[5] = (float)(r[21] * icp);
        float* w = o + FAR_IN + FAR_OUT;
        w[0] = s->grav && sg0 > 0 ? (float)((p->mass * sg0) * (p->mass * sg0) / f2ref) : 0.f;
        w[1] = s->coul && sc0 > 0 ? (float)((p->charge * sc0) * (p->charge * sc0) / f2ref) : 0.f;
        w[2] = s->grav && sp0 > 0 ? (float)((p->mass * sp0) * (p->mass * sp0) / pe2ref) : 0.f;
        w[3] = s->coul && cp0 > 0 ? (float)((p->charge * cp0) * (p->charge * cp0) / pe2ref) : 0.f;
        if (obs) obs[k] = (int32_t)r[0];
    }
    free(raw);
    return m;
}

static void resetForces(OrcSys* s) {                                      /* forces.js:31-54 */
    for (int i = 0; i < s->n; i++) {
        P* p = &s->p[i];
        p->fx = p->fy = 0;
        p->f1x = p->f1y = 0;
        p->Bz = p->dBzdx = p->dBzdy = 0;
        p->Bgz = p->dBgzdx = p->dBgzdy = 0;
        p->frameDrag = p->tidal = p->contactTorque = 0;
        p->jx = p->jy = 0;
        p->pe = 0;
    }
}

/* _applyBoundaryForces (integrator.js:606-667): Hertz repulsion delta^1.5 of
 * each wall the body overlaps, with a tangential friction force clamp(10 vt)
 * times -friction Fn and its contact torque r Ft; domain [0, W] x [0, H]. */
static void applyBoundaryForces(OrcSys* s) {
    const double friction = s->prm.bounce_friction;
    const double left = 0, top = 0, right = s->prm.domain_w, bottom = s->prm.domain_h;
    for (int i = 0; i < s->n; i++) {
        P* p = &s->p[i];
        const double r = p->radius;
        double delta = r - (p->x - left);
        if (delta > 0) {
            const double Fn = delta * sqrt(delta);
            p->fx += Fn;
            if (friction > 0) {
                const double vt = p->vy + p->angVel * r;
                const double Ft = -friction * Fn * fmax(-1.0, fmin(1.0, vt * 10));
                p->fy += Ft;
                p->contactTorque += r * Ft;
            }
        }
        delta = r - (right - p->x);
        if (delta > 0) {
            const double Fn = delta * sqrt(delta);
            p->fx -= Fn;
            if (friction > 0) {
                const double vt = -p->vy - p->angVel * r;
                const double Ft = -friction * Fn * fmax(-1.0, fmin(1.0, vt * 10));
                p->fy -= Ft;
                p->contactTorque -= r * Ft;
            }
        }
        delta = r - (p->y - top);
        if (delta > 0) {
            const double Fn = delta * sqrt(delta);
            p->fy += Fn;
            if (friction > 0) {
                const double vt = -p->vx + p->angVel * r;
                const double Ft = -friction * Fn * fmax(-1.0, fmin(1.0, vt * 10));
                p->fx -= Ft;
                p->contactTorque += r * Ft;
            }
        }
        delta = r - (bottom - p->y);
        if (delta > 0) {
            const double Fn = delta * sqrt(delta);
            p->fy -= Fn;
            if (friction > 0) {
                const double vt = p->vx - p->angVel * r;
                const double Ft = -friction * Fn * fmax(-1.0, fmin(1.0, vt * 10));
                p->fx += Ft;
                p->contactTorque -= r * Ft;
            }
        }
    }
}

/* computeAllForces (forces.js:61-154). */
static void computeAllForces(OrcSys* s) {
    const int n = s->n;
    for (int i = 0; i < n; i++) {                                         /* forces.js:71-76 */
        P* p = &s->p[i];
        p->magMoment = MAG_MOMENT_K * p->charge * p->angVel * p->bodyRadiusSq;
        p->angMomentum = INERTIA_K * p->mass * p->angVel * p->bodyRadiusSq;
    }
    int64_t c0 = 0, c1 = 0, c2 = 0;
    if (s->barnesHut) {
        if (s->root >= 0) {
            #pragma omp parallel num_threads(g_threads) reduction(+:c0,c1,c2)
            {
                int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->pool.count * 2 + 16));
                int64_t cnt[3] = {0, 0, 0};
                #pragma omp for schedule(dynamic, 64)
                for (int i = 0; i < n; i++) calculateForce(s, i, stack, cnt);
                c0 += cnt[0]; c1 += cnt[1]; c2 += cnt[2];
                free(stack);
            }
        }
    } else {
        #pragma omp parallel for num_threads(g_threads) schedule(dynamic, 16) reduction(+:c0)
        for (int i = 0; i < n; i++) {                                     /* forces.js:93-126 */
            P* p = &s->p[i];
            for (int j = 0; j < n; j++) {
                if (i == j) continue;
                const P* o = &s->p[j];
                if (s->delay) {
                    Delayed d;
                    if (!delayedState(s, o, p->x, p->y, &d)) continue;
                    const double r2 = o->bodyRadiusSq;
                    const double av = d.angw / sqrt(1 + d.angw * d.angw * r2);
                    pairForce(s, p, d.x, d.y, d.vx, d.vy, o->mass, o->charge, av, MAG_MOMENT_K * o->charge * av * r2,
                              INERTIA_K * o->mass * av * r2, o->axMod, o->yukMod, o->higgsMod, 1, 1);
                } else {
                    pairForce(s, p, o->x, o->y, o->vx, o->vy, o->mass, o->charge, o->angVel,
                              o->magMoment, o->angMomentum, o->axMod, o->yukMod, o->higgsMod, 1, 0);
                }
                c0++;
            }
        }
    }
    if (s->delay && s->ndead > 0) {                                       /* forces.js:128-153 */
        #pragma omp parallel for num_threads(g_threads) schedule(dynamic, 16) reduction(+:c0)
        for (int i = 0; i < n; i++) {
            int64_t cnt[1] = {0};
            deadForces(s, &s->p[i], cnt);
            c0 += cnt[0];
        }
    }
    s->cnt_pair = c0; s->cnt_node = c1; s->cnt_visit = c2;
}

/* _applyExternalFields (integrator.js:563-592). */
static void applyExternalFields(OrcSys* s) {
    const GeonParams* g = &s->prm;
    if (g->ext_gravity == 0 && g->ext_electric == 0 && g->ext_bz == 0) return;
    const double gx = g->ext_gravity * cos(g->ext_gravity_angle), gy = g->ext_gravity * sin(g->ext_gravity_angle);
    const double ex = g->ext_electric * cos(g->ext_electric_angle), ey = g->ext_electric * sin(g->ext_electric_angle);
    for (int i = 0; i < s->n; i++) {
        P* p = &s->p[i];
        if (g->ext_gravity != 0) { p->fx += p->mass * gx; p->fy += p->mass * gy; }
        if (g->ext_electric != 0) { p->fx += p->charge * ex; p->fy += p->charge * ey; }
        if (g->ext_bz != 0) p->Bz += g->ext_bz;
    }
}

/* _accum1PN (forces.js:445-506): observer state passed explicitly. */
static void accum1PN(const OrcSys* s, P* p, double sx, double sy, double svx, double svy,
                     double sMass, double sCharge, double sYukMod, double sHiggsMod) {
    const GeonParams* g = &s->prm;
    double rx, ry;
    if (s->periodic) minImage(p->x, p->y, sx, sy, g->topology, g->domain_w, g->domain_h, g->domain_w * 0.5, g->domain_h * 0.5, &rx, &ry);
    else { rx = sx - p->x; ry = sy - p->y; }
    const double rSq = rx * rx + ry * ry + s->softeningSq;
    const double invRSq = 1 / rSq;
    const double invR = sqrt(invRSq);
    const double r = 1 / invR;
    const double nx = rx * invR, ny = ry * invR;
    const double pvx = p->vx, pvy = p->vy, pMass = p->mass, pCharge = p->charge;
    if (s->gm) {
        const double v1Sq = pvx * pvx + pvy * pvy, v2Sq = svx * svx + svy * svy;
        const double nDotV1 = nx * pvx + ny * pvy, nDotV2 = nx * svx + ny * svy;
        const double radial = -v1Sq - 2 * v2Sq + 1.5 * nDotV2 * nDotV2 + 5 * pMass * invR + 4 * sMass * invR;
        const double v1Coeff = 4 * nDotV1 - 3 * nDotV2, v2Coeff = 3 * nDotV2;
        const double base = pMass * sMass * invRSq * invR;
        p->f1x += base * (rx * radial + (pvx * v1Coeff + svx * v2Coeff) * r);
        p->f1y += base * (ry * radial + (pvy * v1Coeff + svy * v2Coeff) * r);
    }
    if (s->mag) {
        const double v2DotN = svx * nx + svy * ny, v1DotN = pvx * nx + pvy * ny;
        const double coeff = 0.5 * pCharge * sCharge * invRSq;
        p->f1x += coeff * (pvx * v2DotN - 3 * nx * v1DotN * v2DotN);
        p->f1y += coeff * (pvy * v2DotN - 3 * ny * v1DotN * v2DotN);
    }
    if (s->gm && s->mag) {
        const double crossCoeff = pCharge * sCharge * (pMass + sMass) - (pCharge * pCharge * sMass + sCharge * sCharge * pMass);
        const double fDir = crossCoeff * invRSq * invRSq;
        p->f1x += rx * fDir; p->f1y += ry * fDir;
    }
    if (s->yuk) {
        const double mu = s->higgs ? g->yukawa_mu * sqrt(p->higgsMod * sHiggsMod) : g->yukawa_mu;
        const double expMuR = exp(-mu * r);
        const double nDotV1 = nx * pvx + ny * pvy, nDotV2 = nx * svx + ny * svy;
        const double v1DotV2 = pvx * svx + pvy * svy;
        const double alpha = 1 + mu * r;
        const double beta = 0.5 * g->yukawa_coupling * sqrt(p->yukMod * sYukMod) * pMass * sMass * expMuR * invRSq;
        const double radial = -(alpha * v1DotV2 + (alpha * alpha + alpha + 1) * nDotV1 * nDotV2);
        p->f1x += beta * (radial * nx + alpha * (nDotV2 * pvx + nDotV1 * svx));
        p->f1y += beta * (radial * ny + alpha * (nDotV2 * pvy + nDotV1 * svy));
    }
}

/* compute1PN (forces.js:578-610) and its tree walk (511-571).  Non-ghost leaf
 * sources (and every direct-sum source) act from their past light cone
 * (542-546, 600-602), in the reference whether or not relativity is on;
 * with signal delay OFF they act at the current time. */
static void compute1PN(OrcSys* s) {
    const int n = s->n;
    for (int i = 0; i < n; i++) s->p[i].f1x = s->p[i].f1y = 0;
    if (s->barnesHut) {
        buildTree(s);
        const Pool* q = &s->pool;
        const double thetaSq = s->prm.theta * s->prm.theta;
        const double W = s->prm.domain_w, H = s->prm.domain_h;
        #pragma omp parallel num_threads(g_threads)
        {
            int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)(q->count * 2 + 16));
            #pragma omp for schedule(dynamic, 64)
            for (int pi = 0; pi < n; pi++) {
                P* p = &s->p[pi];
                int top = 0;
                stack[top++] = s->root;
                while (top > 0) {
                    int ni = stack[--top];
                    if (q->mass[ni] < EPS_) continue;
                    double dx, dy;
                    if (s->periodic) minImage(p->x, p->y, q->comX[ni], q->comY[ni], s->prm.topology, W, H, W * 0.5, H * 0.5, &dx, &dy);
                    else { dx = q->comX[ni] - p->x; dy = q->comY[ni] - p->y; }
                    const double dSq = dx * dx + dy * dy, size = q->bw[ni] * 2;
                    if (!q->divided[ni] && q->pointCount[ni] > 0) {
                        size_t base = (size_t)ni * QT_CAP;
                        for (int k = 0; k < q->pointCount[ni]; k++) {
                            const P* o = &s->tp[q->points[base + k]];
                            if (o->original == pi) continue;
                            const P* real = &s->p[o->original];
                            if (s->delay1pn && !o->isGhost) {
                                Delayed d;
                                if (!delayedState(s, real, p->x, p->y, &d)) continue;
                                accum1PN(s, p, d.x, d.y, d.vx, d.vy, o->mass, o->charge, real->yukMod, real->higgsMod);
                            } else {
                                accum1PN(s, p, o->x, o->y, o->vx, o->vy, o->mass, o->charge, real->yukMod, real->higgsMod);
                            }
                        }
                    } else if (q->divided[ni] && size * size < thetaSq * dSq && window_ok(s, q, ni, p->x, p->y, dx, dy)) {
                        const double m = q->mass[ni];
                        accum1PN(s, p, q->comX[ni], q->comY[ni], q->momX[ni] / m, q->momY[ni] / m, m, q->charge[ni], 1, 1);
                    } else if (q->divided[ni]) {
                        stack[top++] = q->nw[ni]; stack[top++] = q->ne[ni];
                        stack[top++] = q->sw[ni]; stack[top++] = q->se[ni];
                    }
                }
            }
            free(stack);
        }
    } else {
        #pragma omp parallel for num_threads(g_threads) schedule(dynamic, 16)
        for (int i = 0; i < n; i++) {
            P* p = &s->p[i];
            for (int j = 0; j < n; j++) {
                if (i == j) continue;
                const P* o = &s->p[j];
                if (s->delay1pn) {
                    Delayed d;
                    if (!delayedState(s, o, p->x, p->y, &d)) continue;
                    accum1PN(s, p, d.x, d.y, d.vx, d.vy, o->mass, o->charge, o->yukMod, o->higgsMod);
                } else {
                    accum1PN(s, p, o->x, o->y, o->vx, o->vy, o->mass, o->charge, o->yukMod, o->higgsMod);
                }
            }
        }
    }
}

#include "geon_oracle_fields.inc"

/* ------------------------------------------------------------------ */
/* public API                                                          */
/* ------------------------------------------------------------------ */
static void syncToggles(OrcSys* s) {                                      /* integrator.js:138-150 */
    const uint32_t t = s->prm.toggles0;
    s->grav = !!(t & GEON_GRAVITY); s->coul = !!(t & GEON_COULOMB);
    s->mag = !!(t & GEON_MAGNETIC); s->gm = !!(t & GEON_GRAVITOMAG);
    s->onePN = !!(t & GEON_ONE_PN); s->rel = !!(t & GEON_RELATIVITY);
    s->spinOrbit = !!(t & GEON_SPIN_ORBIT); s->bh = !!(t & GEON_BLACK_HOLE);
    s->yuk = !!(t & GEON_YUKAWA); s->higgs = !!(t & GEON_HIGGS); s->axion = !!(t & GEON_AXION);
    s->barnesHut = !!(t & GEON_BARNES_HUT); s->expansion = !!(t & GEON_EXPANSION);
    s->rad = !!(t & GEON_RADIATION);
    s->softening = s->bh ? s->prm.bh_softening : s->prm.softening;
    s->softeningSq = s->softening * s->softening;
    s->needAxMod = (s->coul || s->mag) && s->axion;                       /* forces.js:79 */
    const double muMin = s->higgs ? s->prm.yukawa_mu * HIGGS_MASS_FLOOR : s->prm.yukawa_mu;
    s->yukCutoffSq = s->yuk ? (6 / muMin) * (6 / muMin) : INFINITY;       /* forces.js:82-83 */
    s->periodic = s->prm.boundary_mode == GEON_BOUND_LOOP;
    s->delay = s->prm.signal_delay == GEON_DELAY_ON || (s->prm.signal_delay == GEON_DELAY_FOLLOW && s->rel);
    /* compute1PN takes every non-ghost leaf from the light cone whether or not
     * relativity is on (forces.js:542-546, 600-602); without relativity no
     * history is recorded (integrator.js:1549), so those leaves add nothing. */
    s->delay1pn = s->delay || s->prm.signal_delay == GEON_DELAY_FOLLOW;
}

OrcSys* orc_create(const GeonParams* prm, int32_t capacity) {
    OrcSys* s = (OrcSys*)calloc(1, sizeof(OrcSys));
    s->prm = *prm;
    s->capacity = capacity > 0 ? capacity : 1;
    s->p = (P*)calloc((size_t)s->capacity, sizeof(P));
    s->root = -1;
    syncToggles(s);
    const int G = prm->field_grid > 0 ? prm->field_grid : 64;
    of_init(&s->higgsF, G, 1, prm->higgs_mass);
    of_init(&s->axionF, G, 0, prm->axion_mass);
    return s;
}

void orc_destroy(OrcSys* s) {
    if (!s) return;
    of_free(&s->higgsF);
    of_free(&s->axionF);
    pool_free(&s->pool);
    free(s->farModel);
    for (int i = 0; i < s->n; i++) free(s->p[i].hist);
    dead_free(s);
    free(s->dead);
    free(s->ev);
    free(s->p); free(s->tp); free(s);
}

int orc_set_params(OrcSys* s, const GeonParams* prm) { s->prm = *prm; syncToggles(s); return 0; }

/* addParticle normalization (main.js:504-539) with proper velocity given
 * directly (the GPU contract passes w, gpu-physics.js:1857). */
int orc_upload(OrcSys* s, int32_t n, const double* x, const double* y, const double* wx,
               const double* wy, const double* mass, const double* charge, const double* angw) {
    for (int i = 0; i < s->n; i++) free(s->p[i].hist);
    dead_free(s);
    s->histStride = 0;
    if (n > s->capacity) {
        s->capacity = n;
        s->p = (P*)realloc(s->p, sizeof(P) * (size_t)n);
    }
    s->n = n;
    for (int i = 0; i < n; i++) {
        P* p = &s->p[i];
        memset(p, 0, sizeof(*p));
        p->x = x[i]; p->y = y[i];
        p->mass = mass[i]; p->baseMass = mass[i];
        p->q_quanta = (int32_t)lround(charge[i] / BOSON_CHARGE);
        p->charge = (double)p->q_quanta * BOSON_CHARGE;
        p->radius = cbrt(p->mass);
        p->radiusSq = p->radius * p->radius;
        p->bodyRadiusSq = p->radiusSq;
        p->invMass = 1 / p->mass;
        p->angw = angw ? angw[i] : 0;
        p->wx = wx[i]; p->wy = wy[i];
        if (s->rel) {
            const double invG = 1 / sqrt(1 + p->wx * p->wx + p->wy * p->wy);
            p->vx = p->wx * invG; p->vy = p->wy * invG;
        } else { p->vx = p->wx; p->vy = p->wy; }
        p->angVel = s->rel ? angwToAngVel(p->angw, p->radius) : p->angw;
        p->axMod = p->yukMod = p->higgsMod = 1;
        p->id = i;
        p->creationTime = -INFINITY;                /* loaded particles always existed (save-load.js:231) */
        p->deathTime = INFINITY;
    }
    s->next_id = n;
    s->nev = 0;
    s->totalRadiated = s->totalRadiatedPx = s->totalRadiatedPy = 0;
    s->forcesInit = 0;
    s->simTime = 0;
    s->root = -1;
    return 0;
}

int32_t orc_count(OrcSys* s) { return s->n; }

int orc_download(OrcSys* s, double* x, double* y, double* wx, double* wy, double* mass, double* angw, int32_t* ids) {
    for (int i = 0; i < s->n; i++) {
        const P* p = &s->p[i];
        if (x) x[i] = p->x; if (y) y[i] = p->y;
        if (wx) wx[i] = p->wx; if (wy) wy[i] = p->wy;
        if (mass) mass[i] = p->mass; if (angw) angw[i] = p->angw;
        if (ids) ids[i] = p->id;
    }
    return 0;
}

/* Tree rebuild (integrator.js:1340-1348) then resetForces, _syncAxionField and
 * computeAllForces (1490-1493).  The tree's μ and L aggregates use the moments
 * cached by the previous force pass: computeAllForces re-caches them only after
 * the tree is built (forces.js:71-76).  reuseTree: the tree compute1PN built at
 * these positions stands, its aggregates from before the 1PN kick, refreshed
 * only when Higgs changed the masses (integrator.js:1341-1344). */
/* _repelPair (integrator.js:687-724): Hertz contact delta^1.5 along the line of
 * centres, p2 taken where the tree holds it (an image for a ghost), and the
 * friction of the tangential velocity with its contact torques. */
static void repelPair(P* p1, double p2x, double p2y, P* p2, double friction) {
    const double dx = p2x - p1->x, dy = p2y - p1->y;
    const double distSq = dx * dx + dy * dy;
    const double minDist = p1->radius + p2->radius;
    if (distSq >= minDist * minDist) return;
    const double dist = sqrt(distSq);
    const double delta = minDist - dist;
    if (delta <= 0) return;
    const double Fn = delta * sqrt(delta);
    const double safeDist = dist > EPS_ ? dist : EPS_;
    const double nx = dx / safeDist, ny = dy / safeDist;
    p1->fx -= nx * Fn; p1->fy -= ny * Fn;
    p2->fx += nx * Fn; p2->fy += ny * Fn;
    if (friction > 0) {
        const double tx = -ny, ty = nx;
        const double v1t = p1->vx * tx + p1->vy * ty + p1->angVel * p1->radius;
        const double v2t = p2->vx * tx + p2->vy * ty - p2->angVel * p2->radius;
        const double vRel = v1t - v2t;
        const double Ft = -friction * Fn * fmax(-1.0, fmin(1.0, vRel * 10));
        p1->fx += tx * Ft; p1->fy += ty * Ft;
        p2->fx -= tx * Ft; p2->fy -= ty * Ft;
        p1->contactTorque += p1->radius * Ft;
        p2->contactTorque -= p2->radius * Ft;
    }
}

/* _applyRepulsion (integrator.js:669-685): every body queries the substep's tree
 * (ghost images included when periodic) with a box of half-size 2 r and repels
 * each candidate of larger id, in the order the query meets them. */
static void applyRepulsion(OrcSys* s) {
    const Pool* q = &s->pool;
    const double friction = s->prm.bounce_friction;
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)(q->count + 8));
    for (int ci = 0; ci < s->n; ci++) {
        P* p1 = &s->p[ci];
        const double rx = p1->x, ry = p1->y, rw = p1->radius * 2, rh = p1->radius * 2;
        int top = 0;
        stack[top++] = s->root;
        while (top > 0) {
            const int ni = stack[--top];
            if (rx - rw > q->bx[ni] + q->bw[ni] || rx + rw < q->bx[ni] - q->bw[ni] ||
                ry - rh > q->by[ni] + q->bh[ni] || ry + rh < q->by[ni] - q->bh[ni]) continue;
            if (!q->divided[ni]) {
                const size_t base = (size_t)ni * QT_CAP;
                for (int k = 0; k < q->pointCount[ni]; k++) {
                    const P* p2 = &s->tp[q->points[base + k]];
                    if (!(p2->x >= rx - rw && p2->x <= rx + rw && p2->y >= ry - rh && p2->y <= ry + rh)) continue;
                    P* real2 = &s->p[p2->original];
                    if (p2->original == ci || p1->id >= real2->id) continue;
                    repelPair(p1, p2->x, p2->y, real2, friction);
                }
            } else {
                stack[top++] = q->nw[ni]; stack[top++] = q->ne[ni];
                stack[top++] = q->sw[ni]; stack[top++] = q->se[ni];
            }
        }
    }
    free(stack);
}

static void forcePass(OrcSys* s, int reuseTree) {
    if (s->barnesHut) {
        if (!reuseTree) buildTree(s);
        else if (s->higgs) {
            for (int i = 0; i < s->ntp; i++) {
                if (s->tp[i].isGhost) continue;
                const int orig = s->tp[i].original;
                s->tp[i] = s->p[orig];
                s->tp[i].isGhost = 0; s->tp[i].original = orig;
            }
            pool_massDistribution(&s->pool, s->tp, s->root);
            pool_moments(&s->pool, s->tp, s->root);
        }
    }
    resetForces(s);
    if (s->axion) of_axmod(s, &s->axionF);                               /* _syncAxionField (integrator.js:595-605) */
    else for (int i = 0; i < s->n; i++) s->p[i].axMod = s->p[i].yukMod = 1;
    computeAllForces(s);
    applyExternalFields(s);
    if (s->higgs) of_apply_forces(s, &s->higgsF);                        /* integrator.js:1494-1507 */
    if (s->axion) of_apply_forces(s, &s->axionF);
    if (s->grav) {
        if (s->higgs) of_grav_forces(s, &s->higgsF);
        if (s->axion) of_grav_forces(s, &s->axionF);
    }
    if (s->prm.collision_mode == GEON_COL_BOUNCE) {                      /* integrator.js:779, 1545 */
        /* the substep's tree, which the reference always builds, with its ghosts when periodic */
        const int ghostTree = s->periodic && s->prm.periodic_mode == GEON_PERIODIC_GHOSTS;
        if (!s->barnesHut || (s->periodic && !ghostTree)) buildTreeG(s, s->periodic);
        applyRepulsion(s);
    }
    if (s->prm.boundary_mode == GEON_BOUND_BOUNCE) applyBoundaryForces(s);   /* integrator.js:780, 1546 */
}

int orc_compute_forces(OrcSys* s) {
    syncToggles(s);
    for (int i = 0; i < s->n; i++)
        s->p[i].angVel = s->rel ? angwToAngVel(s->p[i].angw, s->p[i].radius) : s->p[i].angw;
    forcePass(s, 0);
    s->forcesInit = 1;
    return 0;
}

int orc_get_forces(OrcSys* s, double* fx, double* fy, double* bz, double* bgz, double* dbzdx,
                   double* dbzdy, double* dbgzdx, double* dbgzdy, double* frame_drag, double* tidal,
                   double* pe_total) {
    double pe = 0;
    for (int i = 0; i < s->n; i++) {
        const P* p = &s->p[i];
        if (fx) fx[i] = p->fx; if (fy) fy[i] = p->fy;
        if (bz) bz[i] = p->Bz; if (bgz) bgz[i] = p->Bgz;
        if (dbzdx) dbzdx[i] = p->dBzdx; if (dbzdy) dbzdy[i] = p->dBzdy;
        if (dbgzdx) dbgzdx[i] = p->dBgzdx; if (dbgzdy) dbgzdy[i] = p->dBgzdy;
        if (frame_drag) frame_drag[i] = p->frameDrag; if (tidal) tidal[i] = p->tidal;
        pe += p->pe;
    }
    if (pe_total) *pe_total = pe * 0.5;                                   /* getPEAccum (forces.js:26) */
    return 0;
}

static void cyclotronScan(OrcSys* s, double* maxCyc) {                    /* integrator.js:817-849 */
    const int hasMagnetic = s->mag, hasExtBz = s->prm.ext_bz != 0, hasGM = s->gm;
    for (int i = 0; i < s->n; i++) {
        const P* p = &s->p[i];
        if (hasMagnetic || hasExtBz) {
            const double absBz = p->Bz > 0 ? p->Bz : -p->Bz;
            if (absBz > 0) {
                const double absQ = p->charge > 0 ? p->charge : -p->charge;
                const double wc = absQ * absBz * p->invMass;
                if (wc > *maxCyc) *maxCyc = wc;
            }
        }
        if (hasGM) {
            const double absBgz = p->Bgz > 0 ? p->Bgz : -p->Bgz;
            if (absBgz > 0) { const double wc = 4 * absBgz; if (wc > *maxCyc) *maxCyc = wc; }
        }
    }
}

/* Physics.update (integrator.js:726-1825), the parts on the hot path. */
#include "geon_oracle_merge.inc"
#include "geon_oracle_rad.inc"

int orc_update(OrcSys* s, double dt, int32_t* substeps) {
    syncToggles(s);
    const GeonParams* g = &s->prm;
    int n = s->n;                  /* particles.length, refreshed after collisions (integrator.js:1386) */
    s->nev = 0;
    const int relOn = s->rel;
    if (!s->forcesInit && n > 0) {                                        /* integrator.js:750-782 */
        for (int i = 0; i < n; i++) s->p[i].angVel = relOn ? angwToAngVel(s->p[i].angw, s->p[i].radius) : s->p[i].angw;
        forcePass(s, 0);
        s->forcesInit = 1;
    }
    const int hasGrav = s->grav, hasMagnetic = s->mag, hasGM = s->gm, hasExtBz = g->ext_bz != 0;
    double dtRemain = dt;
    int totalSteps = 0;
    double prevMaxAccelSq = 0;
    const int maxSub = g->max_substeps;
    while (dtRemain > EPS_ && totalSteps < maxSub) {                      /* integrator.js:810 */
        double maxAccelSq = 0, maxCyc = 0;
        if (totalSteps == 0) {
            for (int i = 0; i < n; i++) {
                const P* p = &s->p[i];
                const double aSq = (p->fx * p->fx + p->fy * p->fy) * p->invMass * p->invMass;
                if (aSq > maxAccelSq) maxAccelSq = aSq;
            }
            cyclotronScan(s, &maxCyc);
        } else {
            maxAccelSq = prevMaxAccelSq;
            if (hasMagnetic || hasExtBz || hasGM) cyclotronScan(s, &maxCyc);
        }
        double aMax = sqrt(maxAccelSq);
        if (aMax != aMax) aMax = 0;
        double dtSafe = aMax > EPS_ ? sqrt(s->softening / aMax) : dtRemain;   /* integrator.js:864 */
        if (maxCyc > 0) {
            const double dtC = 0.7853981633974483 / maxCyc;
            if (dtC < dtSafe) dtSafe = dtC;
        }
        const int budgetLeft = maxSub - totalSteps;
        double need = ceil(dtRemain / dtSafe);
        const int stepsNeeded = need < budgetLeft ? (int)need : budgetLeft;
        const double dtSub = dtRemain / stepsNeeded;
        totalSteps++;
        dtRemain -= dtSub;
        const int has1PN = s->onePN;
        const double halfDt = dtSub * 0.5;
        const int needBoris = hasMagnetic || hasGM || hasExtBz;
        double nextMaxAccelSq = 0;
        for (int i = 0; i < n; i++) {                                     /* integrator.js:884-928 */
            P* p = &s->p[i];
            const double invM = p->invMass, halfDtOverM = halfDt * invM;
            if (has1PN) { p->f1ox = p->f1x; p->f1oy = p->f1y; }
            const double fx = p->fx, fy = p->fy;
            p->wx += fx * halfDtOverM; p->wy += fy * halfDtOverM;
            if (needBoris) {
                double t = 0;
                if (hasMagnetic || hasExtBz) t += (p->charge * 0.5 * invM) * p->Bz;
                if (hasGM) t += 2 * p->Bgz;
                if (t != 0) {
                    const double gamma = relOn ? sqrt(1 + p->wx * p->wx + p->wy * p->wy) : 1;
                    t *= dtSub / gamma;
                    const double sB = 2 * t / (1 + t * t);
                    const double wx = p->wx, wy = p->wy;
                    const double wpx = wx + wy * t, wpy = wy - wx * t;
                    p->wx = wx + wpy * sB; p->wy = wy - wpx * sB;
                }
            }
            p->wx += fx * halfDtOverM; p->wy += fy * halfDtOverM;
            if (p->wx != p->wx || p->wy != p->wy) { p->wx = p->wy = 0; p->vx = p->vy = 0; }
            const double aSq = (fx * fx + fy * fy) * invM * invM;
            if (aSq > nextMaxAccelSq) nextMaxAccelSq = aSq;
        }
        prevMaxAccelSq = nextMaxAccelSq;
        const int needSpinOrbit = s->spinOrbit && (hasMagnetic || hasGM);     /* integrator.js:930-981 */
        const int needTorques = (hasGM && relOn) || hasGrav || s->prm.collision_mode == GEON_COL_BOUNCE;
        if (needSpinOrbit || needTorques) {
            for (int i = 0; i < n; i++) {
                P* p = &s->p[i];
                if (needSpinOrbit && fabs(p->angVel) >= EPS_) {
                    const double IxOmega = p->angMomentum;
                    if (fabs(IxOmega) >= EPS_) {
                        const double dtOverM = dtSub * p->invMass;
                        if (hasMagnetic && fabs(p->charge) > EPS_) {
                            const double mu = p->magMoment;
                            p->angw -= mu * (p->vx * p->dBzdx + p->vy * p->dBzdy) * dtSub / IxOmega;
                            p->wx += mu * p->dBzdx * dtOverM; p->wy += mu * p->dBzdy * dtOverM;
                        }
                        if (hasGM) {
                            const double L = p->angMomentum;
                            p->angw -= L * (p->vx * p->dBgzdx + p->vy * p->dBgzdy) * dtSub / IxOmega;
                            p->wx -= L * p->dBgzdx * dtOverM; p->wy -= L * p->dBgzdy * dtOverM;
                        }
                        if (p->angw != p->angw) p->angw = 0;
                        const double sr = p->angw * p->radius;
                        p->angVel = relOn ? p->angw / sqrt(1 + sr * sr) : p->angw;
                    }
                }
                if (needTorques) {
                    double torque = 0;
                    if (hasGM && relOn) torque += p->frameDrag;
                    if (hasGrav) torque += p->tidal;
                    torque += p->contactTorque;
                    if (torque != 0) {
                        const double I = INERTIA_K * p->mass * p->radiusSq;
                        if (I >= EPS_) {
                            p->angw += torque * dtSub / I;
                            if (p->angw != p->angw) p->angw = 0;
                            const double sr = p->angw * p->radius;
                            p->angVel = relOn ? p->angw / sqrt(1 + sr * sr) : p->angw;
                        }
                    }
                }
            }
        }
        if (s->rad && s->coul) llRadiation(s, dtSub);                    /* integrator.js:983-1086 */
        for (int i = 0; i < n; i++) {                                     /* integrator.js:1270-1278 */
            P* p = &s->p[i];
            const double invG = relOn ? 1 / sqrt(1 + p->wx * p->wx + p->wy * p->wy) : 1;
            p->vx = p->wx * invG; p->vy = p->wy * invG;
            p->angVel = relOn ? angwToAngVel(p->angw, p->radius) : p->angw;
            p->x += p->vx * dtSub; p->y += p->vy * dtSub;
        }
        if (s->expansion) {                                               /* integrator.js:1281-1292 */
            const double Hh = g->hubble, cx = g->domain_w * 0.5, cy = g->domain_h * 0.5;
            for (int i = 0; i < n; i++) {
                P* p = &s->p[i];
                p->x += Hh * (p->x - cx) * dtSub; p->y += Hh * (p->y - cy) * dtSub;
                const double decay = 1 - Hh * dtSub;
                p->wx *= decay; p->wy *= decay;
            }
        }
        s->simTime += dtSub;
        if (has1PN) {                                                     /* integrator.js:1296-1318 */
            if (relOn && s->expansion) {
                for (int i = 0; i < n; i++) {
                    P* p = &s->p[i];
                    const double invG = 1 / sqrt(1 + p->wx * p->wx + p->wy * p->wy);
                    p->vx = p->wx * invG; p->vy = p->wy * invG;
                }
            }
            compute1PN(s);
            for (int i = 0; i < n; i++) {
                P* p = &s->p[i];
                const double halfDtOverM = halfDt * p->invMass;
                p->wx += (p->f1x - p->f1ox) * halfDtOverM;
                p->wy += (p->f1y - p->f1oy) * halfDtOverM;
                if (p->wx != p->wx || p->wy != p->wy) { p->wx = p->wy = 0; p->vx = p->vy = 0; }
            }
        }
        if (s->higgs) {                                                   /* integrator.js:1320-1329 */
            of_update(s, &s->higgsF, dtSub, s->axion ? &s->axionF : NULL);
            of_modulate_masses(s, &s->higgsF, dtSub);
        }
        if (s->axion) of_update(s, &s->axionF, dtSub, s->higgs ? &s->higgsF : NULL);   /* 1331-1334 */
        int reuse = has1PN && s->barnesHut;
        if (g->collision_mode == GEON_COL_MERGE && n > 0) {              /* integrator.js:1336-1416 */
            /* the query finds neighbours across a periodic edge through the
             * ghosts of the step-5 tree (integrator.js:247-258), also where the
             * force walk runs without them */
            const int ghostQuery = s->periodic && g->periodic_mode != GEON_PERIODIC_GHOSTS;
            if (!reuse || ghostQuery) buildTreeG(s, s->periodic);
            if (ghostQuery) reuse = 0;
            /* After a change the next force pass walks a fresh tree: the
             * reference walks the step-5 tree, where removed bodies still sit
             * (mass 0) and spawned ones are missing, for one substep. */
            if (mergeCollisions(s)) reuse = 0;
            n = s->n;
        }
        forcePass(s, reuse);                                              /* integrator.js:1489-1493 */
    }
    if (s->delay && n > 0) {                                              /* integrator.js:1549-1565 */
        const int stride = g->history_stride > 0 ? g->history_stride : 64;
        if (++s->histStride >= stride) {
            s->histStride -= stride;
            for (int i = 0; i < n; i++) hist_push(&s->p[i], s->simTime);
        }
    }
    if (s->rad && n >= 2 && (s->grav || s->coul)) quadRadiation(s, dt);   /* integrator.js:1567-1740 */
    s->potentialEnergy = 0;
    for (int i = 0; i < n; i++) s->potentialEnergy += s->p[i].pe;
    s->potentialEnergy *= 0.5;                                            /* integrator.js:1743 */
    if (s->grav) {                                                        /* integrator.js:1744-1749 */
        if (s->higgs) s->potentialEnergy += of_grav_pe(s, &s->higgsF);
        if (s->axion) s->potentialEnergy += of_grav_pe(s, &s->axionF);
    }
    /* Boundary handling once per update (integrator.js:1798-1824). */
    int w = 0;
    const double left = 0, top = 0, right = g->domain_w, bottom = g->domain_h;
    for (int i = 0; i < s->n; i++) {
        P* p = &s->p[i];
        if (g->boundary_mode == GEON_BOUND_DESPAWN) {
            if (p->x < left - DESPAWN_MARGIN || p->x > right + DESPAWN_MARGIN ||
                p->y < top - DESPAWN_MARGIN || p->y > bottom + DESPAWN_MARGIN) {
                retireParticle(s, p);
                continue;
            }
        } else if (g->boundary_mode == GEON_BOUND_LOOP) {
            wrapPosition(p, g->topology, g->domain_w, g->domain_h);
        } else if (g->boundary_mode == GEON_BOUND_BOUNCE) {
            if (p->x < left) p->x = left; else if (p->x > right) p->x = right;
            if (p->y < top) p->y = top; else if (p->y > bottom) p->y = bottom;
        }
        if (w != i) s->p[w] = *p;
        w++;
    }
    s->n = w;
    dead_gc(s);                 /* once per update here; once per frame in main.js:874-885 */
    if (substeps) *substeps = totalSteps;
    return 0;
}

/* ---- signal-delay state for tests ---- */
int32_t orc_get_history(OrcSys* s, int32_t i, int32_t dead, double* x, double* y, double* vx, double* vy,
                        double* aw, double* t) {
    const P* p = dead ? (i < s->ndead ? &s->dead[i] : NULL) : (i < s->n ? &s->p[i] : NULL);
    if (!p || !p->hist) return 0;
    const Hist* h = p->hist;
    const int start = (h->head - h->count + HIST_N) & HIST_MASK;
    for (int k = 0; k < h->count; k++) {
        const int j = (start + k) & HIST_MASK;
        x[k] = h->x[j]; y[k] = h->y[j]; vx[k] = h->vx[j]; vy[k] = h->vy[j]; aw[k] = h->aw[j]; t[k] = h->t[j];
    }
    return h->count;
}

int orc_set_history(OrcSys* s, int32_t i, int32_t count, const double* x, const double* y, const double* vx,
                    const double* vy, const double* aw, const double* t) {
    if (i < 0 || i >= s->n || count < 0 || count > HIST_N) return -1;
    P* p = &s->p[i];
    hist_init(p);
    memset(p->hist, 0, sizeof(Hist));
    for (int k = 0; k < count; k++) {
        Hist* h = p->hist;
        h->x[k] = x[k]; h->y[k] = y[k]; h->vx[k] = vx[k]; h->vy[k] = vy[k]; h->aw[k] = aw[k]; h->t[k] = t[k];
    }
    p->hist->count = count;
    p->hist->head = count & HIST_MASK;
    return 0;
}

int orc_get_f1pn(OrcSys* s, double* f1x, double* f1y) {
    for (int i = 0; i < s->n; i++) { f1x[i] = s->p[i].f1x; f1y[i] = s->p[i].f1y; }
    return 0;
}

int orc_min_image(OrcSys* s, double ox, double oy, double sx, double sy, double* out2) {
    syncToggles(s);
    const GeonParams* g = &s->prm;
    if (s->periodic) minImage(ox, oy, sx, sy, g->topology, g->domain_w, g->domain_h, g->domain_w * 0.5, g->domain_h * 0.5, &out2[0], &out2[1]);
    else { out2[0] = sx - ox; out2[1] = sy - oy; }
    return 0;
}

void orc_set_time(OrcSys* s, double t) { s->simTime = t; }
int32_t orc_dead_count(OrcSys* s) { return s->ndead; }
int orc_set_lifetime(OrcSys* s, int32_t i, double creation, double death) {
    if (i < 0 || i >= s->n) return -1;
    s->p[i].creationTime = creation;
    s->p[i].deathTime = death;
    return 0;
}

int orc_delayed_state(OrcSys* s, int32_t i, double ox, double oy, double* out5) {
    syncToggles(s);
    if (i < 0 || i >= s->n) return -1;
    Delayed d;
    if (!delayedState(s, &s->p[i], ox, oy, &d)) return 0;
    out5[0] = d.x; out5[1] = d.y; out5[2] = d.vx; out5[3] = d.vy; out5[4] = d.angw;
    return 1;
}

/* computeEnergies (energy.js:11-195), particle passes 1-3; PE from the last
 * force pass (integrator.js:1743). */
int orc_stats(OrcSys* s, GeonStats* out) {
    memset(out, 0, sizeof(*out));
    out->radiated_energy = s->totalRadiated;
    out->radiated_px = s->totalRadiatedPx;
    out->radiated_py = s->totalRadiatedPy;
    const int n = s->n, relOn = s->rel;
    double linearKE = 0, spinKE = 0, totalMass = 0, comX = 0, comY = 0, px = 0, py = 0;
    for (int i = 0; i < n; i++) {
        const P* p = &s->p[i];
        const double pm = p->mass, rSq = p->radiusSq;
        if (relOn) {
            const double wSq = p->wx * p->wx + p->wy * p->wy;
            const double gamma = sqrt(1 + wSq);
            linearKE += wSq / (gamma + 1) * pm;
            const double srSq = p->angw * p->angw * rSq;
            const double gammaRot = sqrt(1 + srSq);
            spinKE += INERTIA_K * pm * srSq / (gammaRot + 1);
        } else {
            const double speedSq = p->vx * p->vx + p->vy * p->vy;
            linearKE += 0.5 * pm * speedSq;
            spinKE += 0.5 * INERTIA_K * pm * rSq * p->angVel * p->angVel;
        }
        px += pm * p->wx; py += pm * p->wy;
        totalMass += pm; comX += pm * p->x; comY += pm * p->y;
    }
    double orb = 0, spin = 0;
    if (totalMass > 0) {
        comX /= totalMass; comY /= totalMass;
        for (int i = 0; i < n; i++) {
            const P* p = &s->p[i];
            const double dx = p->x - comX, dy = p->y - comY;
            orb += dx * (p->mass * p->wy) - dy * (p->mass * p->wx);
            spin += INERTIA_K * p->mass * p->radiusSq * p->angw;
        }
    }
    double fieldEnergy = 0, fieldPx = 0, fieldPy = 0;
    if (s->mag || s->gm) {                                                /* energy.js:64-142 */
        const int emOn = s->mag && !s->onePN, gmOn = s->gm && !s->onePN;
        const double W = s->prm.domain_w, H = s->prm.domain_h;
        for (int i = 0; i < n; i++) {
            const P* pi = &s->p[i];
            for (int j = i + 1; j < n; j++) {
                const P* pj = &s->p[j];
                double sx = pj->x, sy = pj->y, svx = pj->vx, svy = pj->vy;
                if (s->delay) {                                           /* energy.js:85-93 */
                    Delayed d;
                    if (!delayedState(s, pj, pi->x, pi->y, &d)) continue;
                    sx = d.x; sy = d.y; svx = d.vx; svy = d.vy;
                }
                double dx, dy;
                if (s->periodic) minImage(pi->x, pi->y, sx, sy, s->prm.topology, W, H, W * 0.5, H * 0.5, &dx, &dy);
                else { dx = sx - pi->x; dy = sy - pi->y; }
                const double rSq = dx * dx + dy * dy + s->softeningSq;
                const double invR = 1 / sqrt(rSq);
                const double rx = dx * invR, ry = dy * invR;
                const double viDotVj = pi->vx * svx + pi->vy * svy;
                const double viDotR = pi->vx * rx + pi->vy * ry;
                const double vjDotR = svx * rx + svy * ry;
                const double velTerm = viDotVj + viDotR * vjDotR;
                const double sumVx = pi->vx + svx, sumVy = pi->vy + svy;
                const double svDotR = sumVx * rx + sumVy * ry;
                if (s->mag) {
                    const double qqInvR = pi->charge * pj->charge * invR * sqrt(pi->axMod * pj->axMod);
                    if (emOn) fieldEnergy -= 0.5 * qqInvR * velTerm;
                    const double coeff = qqInvR * 0.5;
                    fieldPx += coeff * (sumVx + rx * svDotR);
                    fieldPy += coeff * (sumVy + ry * svDotR);
                }
                if (s->gm) {
                    const double mmInvR = pi->mass * pj->mass * invR;
                    if (gmOn) fieldEnergy += 0.5 * mmInvR * velTerm;
                    const double coeff = mmInvR * 0.5;
                    fieldPx -= coeff * (sumVx + rx * svDotR);
                    fieldPy -= coeff * (sumVy + ry * svDotR);
                }
                if (s->mag && s->gm && !s->onePN) {
                    const double invRSq = invR * invR;
                    const double cc = pi->charge * pj->charge * (pi->mass + pj->mass)
                        - (pi->charge * pi->charge * pj->mass + pj->charge * pj->charge * pi->mass);
                    fieldEnergy += 0.5 * cc * invRSq;
                }
            }
        }
    }
    out->linear_ke = linearKE; out->spin_ke = spinKE; out->px = px; out->py = py;
    out->orbital_ang_mom = orb; out->spin_ang_mom = spin;
    out->com_x = comX; out->com_y = comY; out->total_mass = totalMass;
    out->pe = s->potentialEnergy;
    out->field_energy = fieldEnergy; out->field_px = fieldPx; out->field_py = fieldPy;
    const double W = s->prm.domain_w, H = s->prm.domain_h;
    if (s->higgs) {                                                       /* energy.js:146-176 */
        out->higgs_field_energy = of_field_energy(&s->higgsF, W, H);
        if (s->axion) {                                                   /* higgs-field.js:353-364 */
            const int G = s->higgsF.G;
            double t = 0;
            for (int i = 0; i < G * G; i++) t += s->higgsF.field[i] * s->higgsF.field[i] * s->axionF.field[i] * s->axionF.field[i];
            out->higgs_field_energy += 0.5 * HIGGS_AXION_COUPLING * t * (W / G) * (H / G);
        }
        out->pfi_energy += of_pfi_energy(s, &s->higgsF);
    }
    if (s->axion) {
        out->axion_field_energy = of_field_energy(&s->axionF, W, H);
        out->pfi_energy += of_pfi_energy(s, &s->axionF);
    }
    for (int w = 0; w < 2; w++) {                                         /* energy.js:176-184 */
        if (!(w == 0 ? s->higgs : s->axion)) continue;
        double mx, my;
        of_momentum(w == 0 ? &s->higgsF : &s->axionF, W, H, &mx, &my);
        out->field_px += mx;
        out->field_py += my;
    }
    out->pe += out->pfi_energy;                                           /* energy.js:186 */
    out->sim_time = s->simTime;
    out->alive_count = n;
    return 0;
}

int64_t orc_build_tree(OrcSys* s) {
    syncToggles(s);
    buildTree(s);
    return s->pool.count;
}

int64_t orc_get_tree(OrcSys* s, GeonTreeNode* out, int64_t cap) {
    const Pool* q = &s->pool;
    int64_t m = q->count < cap ? q->count : cap;
    for (int64_t i = 0; i < m; i++) {
        out[i].path = q->path[i];
        out[i].level = q->level[i];
        out[i].divided = q->divided[i];
        out[i].count = q->subCount[i];
        out[i].charge_quanta = q->qq[i];
        out[i].mass = q->mass[i];
        out[i].com_x = q->comX[i];
        out[i].com_y = q->comY[i];
    }
    return q->count;
}

void orc_counters(OrcSys* s, int64_t* out3) { out3[0] = s->cnt_pair; out3[1] = s->cnt_node; out3[2] = s->cnt_visit; }

static int cmp_pair(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : x > y;
}

/* Overlap pairs found by tree query (collisions.js:35-60, quadtree.js:372-404). */
int64_t orc_collision_pairs(OrcSys* s, int32_t* a, int32_t* b, int64_t cap) {
    orc_build_tree(s);
    const Pool* q = &s->pool;
    const double W = s->prm.domain_w, H = s->prm.domain_h;
    int64_t np = 0, pcap = 1024;
    int64_t* pairs = (int64_t*)malloc(sizeof(int64_t) * (size_t)pcap);
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)(q->count + 8));
    for (int ci = 0; ci < s->n; ci++) {
        const P* p1 = &s->p[ci];
        if (p1->mass == 0) continue;
        const double rx = p1->x, ry = p1->y, rw = p1->radius * 2, rh = p1->radius * 2;
        int top = 0;
        stack[top++] = s->root;
        while (top > 0) {
            int ni = stack[--top];
            if (rx - rw > q->bx[ni] + q->bw[ni] || rx + rw < q->bx[ni] - q->bw[ni] ||
                ry - rh > q->by[ni] + q->bh[ni] || ry + rh < q->by[ni] - q->bh[ni]) continue;
            if (!q->divided[ni]) {
                size_t base = (size_t)ni * QT_CAP;
                for (int k = 0; k < q->pointCount[ni]; k++) {
                    const P* p2 = &s->tp[q->points[base + k]];
                    if (!(p2->x >= rx - rw && p2->x <= rx + rw && p2->y >= ry - rh && p2->y <= ry + rh)) continue;
                    const P* real2 = &s->p[p2->original];
                    if (p2->original == ci || real2->mass == 0 || p1->id >= real2->id) continue;
                    double dx, dy;
                    if (s->periodic) minImage(p1->x, p1->y, p2->x, p2->y, s->prm.topology, W, H, W * 0.5, H * 0.5, &dx, &dy);
                    else { dx = p2->x - p1->x; dy = p2->y - p1->y; }
                    const double minDist = p1->radius + real2->radius;
                    if (dx * dx + dy * dy < minDist * minDist) {
                        if (np == pcap) { pcap *= 2; pairs = (int64_t*)realloc(pairs, sizeof(int64_t) * (size_t)pcap); }
                        pairs[np++] = ((int64_t)p1->id << 32) | (uint32_t)real2->id;
                    }
                }
            } else {
                stack[top++] = q->nw[ni]; stack[top++] = q->ne[ni];
                stack[top++] = q->sw[ni]; stack[top++] = q->se[ni];
            }
        }
    }
    qsort(pairs, (size_t)np, sizeof(int64_t), cmp_pair);
    for (int64_t k = 0; k < np && k < cap; k++) { a[k] = (int32_t)(pairs[k] >> 32); b[k] = (int32_t)(pairs[k] & 0xffffffff); }
    free(pairs); free(stack);
    return np;
}

/* FP64 direct sum over all sources for selected targets (forces.js:93-126),
 * min-image when periodic.  Leaves s->p force fields of non-target particles
 * untouched; writes totals into fx/fy (indexed like idx). */
int orc_direct_forces(OrcSys* s, const int32_t* idx, int32_t n_idx, double* fx, double* fy) {
    syncToggles(s);
    const int n = s->n;
    for (int i = 0; i < n; i++) {
        P* p = &s->p[i];
        p->magMoment = MAG_MOMENT_K * p->charge * p->angVel * p->bodyRadiusSq;
        p->angMomentum = INERTIA_K * p->mass * p->angVel * p->bodyRadiusSq;
    }
    const int m = idx ? n_idx : n;
    #pragma omp parallel for num_threads(g_threads) schedule(dynamic, 8)
    for (int k = 0; k < m; k++) {
        const int i = idx ? idx[k] : k;
        P tmp = s->p[i];
        tmp.fx = tmp.fy = 0; tmp.f1x = tmp.f1y = 0; tmp.pe = 0;
        tmp.Bz = tmp.dBzdx = tmp.dBzdy = tmp.Bgz = tmp.dBgzdx = tmp.dBgzdy = 0;
        tmp.frameDrag = tmp.tidal = 0;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            const P* o = &s->p[j];
            pairForce(s, &tmp, o->x, o->y, o->vx, o->vy, o->mass, o->charge, o->angVel,
                      o->magMoment, o->angMomentum, o->axMod, o->yukMod, o->higgsMod, 1, 0);
        }
        fx[k] = tmp.fx; fy[k] = tmp.fy;
    }
    return 0;
}

/* ---- scalar-field access for tests ---- */
static OField* of_pick(OrcSys* s, int which) { return which == 0 ? &s->higgsF : &s->axionF; }

int orc_field_set(OrcSys* s, int32_t which, const double* field, const double* field_dot) {
    OField* f = of_pick(s, which);
    const int G = f->G;
    if (field) memcpy(f->field, field, sizeof(double) * G * G);
    if (field_dot) memcpy(f->fieldDot, field_dot, sizeof(double) * G * G);
    of_grid_gradients(f, s->prm.boundary_mode, s->prm.topology);
    return 0;
}

/* what: 0 field, 1 fieldDot, 2 gradX, 3 gradY, 4 sgPhi, 5 sgGx, 6 sgGy, 7 source, 8 thermal, 9 rho */
int orc_field_get(OrcSys* s, int32_t which, int32_t what, double* out) {
    OField* f = of_pick(s, which);
    const double* src[] = {f->field, f->fieldDot, f->gradX, f->gradY, f->sgPhi, f->sgGx, f->sgGy, f->source, f->thermal, f->rho};
    if (what < 0 || what > 9) return -1;
    memcpy(out, src[what], sizeof(double) * f->G * f->G);
    return f->G;
}

int orc_field_selfgrav(OrcSys* s, int32_t which) {
    syncToggles(s);
    OField* f = of_pick(s, which);
    of_selfgrav(f, s->prm.domain_w, s->prm.domain_h, s->softeningSq, s->prm.boundary_mode, s->prm.topology);
    return f->hasEnergy;
}

/* PQS deposit of per-particle values into a G x G grid (scalar-field.js:145-180). */
int orc_pqs_deposit(int32_t G, double W, double H, int32_t bc, int32_t n, const double* x, const double* y,
                    const double* v, double* out) {
    memset(out, 0, sizeof(double) * G * G);
    for (int i = 0; i < n; i++) {
        int ix, iy; double wx[4], wy[4];
        of_pqs(x[i], y[i], G / W, G / H, &ix, &iy, wx, wy);
        of_deposit_raw(G, out, ix, iy, wx, wy, v[i], bc, GEON_TORUS);
    }
    return 0;
}
