/*
 * nddo_oracle.c -- CPU restatement of MolDS's NDDO (MNDO/AM1/PM3) ground-state
 * SCF building blocks, used ONLY as test infrastructure: the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it as the checker.
 * The product (molds_amd/csrc, libnddo.so) never links or calls this file.
 *
 * It follows the reference function by function, in plain C99, FP64 only:
 *   parameters ........ src/base/atoms/{H,C,N,O,F}atom.cpp SetAtomicParameters,
 *                        src/base/atoms/Atom.cpp:263-1109 accessors,
 *                        src/base/Parameters.cpp:45-48 unit constants
 *   geometry .......... src/base/Molecule.cpp:422-438 (distances),
 *                        :512-525 (AO / electron counts)
 *   rotation .......... src/cndo/Cndo2.cpp:5498-5531 (p block, gamma = 0),
 *                        src/base/EularAngle.cpp:38-61
 *   overlap ........... src/cndo/Cndo2.cpp:4030-4146, 5801-5869, 6091-6149,
 *                        6185-6204 (B.24), 6385-6477 (A_k, B_k, D)
 *   multipoles ........ src/mndo/Mndo.cpp:6592-6834 (Dewar-Thiel eqs. 52-66)
 *   (mu nu|lam sig) ... src/mndo/Mndo.cpp:4886-5238 (eqs. 29-49), 3948-4025,
 *                        4223-4293 (rotation), 3703-3820 (8-fold fill)
 *   Fock .............. src/cndo/Cndo2.cpp:1526-1636, src/mndo/Mndo.cpp:422-650
 *   core repulsion .... src/mndo/Mndo.cpp:223-240, 319-327; src/am1/Am1.cpp:145-170,
 *                        src/am1/Am1.h:45-47; src/cndo/Cndo2.cpp:406-416
 *   electronic energy . src/cndo/Cndo2.cpp:1324-1416
 *   heat of formation . src/mndo/Mndo.cpp:395-404 (MNDO only)
 *
 * The reduced-overlap coefficients Y_ij of (B.24) are not read from MolDS's
 * generated table; they are expanded here from the prolate-spheroidal form of
 * the Slater product (see orc_overlap_poly), which the tests pin against the
 * closed form of the 1s-1s overlap and against brute-force 3-D quadrature.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Neumaier's compensated summation: the energies are differences of sums of
 * order 1e5 Hartree over up to millions of terms, which plain summation
 * resolves only to ~1e-12 relative. */
static void nsum_add(double* s, double* c, double x)
{
    const double t = *s + x;
    if (fabs(*s) >= fabs(x)) *c += (*s - t) + x;
    else *c += (x - t) + *s;
    *s = t;
}

#define ORC_EV2AU      0.03674903
#define ORC_KCAL2AU    0.00159360175
static const double ORC_ANG2AU = 1.0 / 0.5291772;

enum { ORC_MNDO = 0, ORC_AM1 = 1, ORC_PM3 = 2 };
/* AO order inside an atom: s, py, pz, px (src/base/Enums.h OrbitalType). */
enum { S = 0, PY = 1, PZ = 2, PX = 3 };

/* ---------------------------------------------------------------- params */
/* Published parameter values, in the units MolDS writes them (eV, 1/Angstrom,
 * Angstrom); converted in orc_elem_params exactly as the reference does. */
typedef struct {
    double Uss, Upp, zs, zp, bs, bp, alpha;
    double Gss, Gpp, Gsp, Gpp2, Hsp;
    double D1, D2, rho0i, rho1i, rho2i; /* rho_k = 0.5/rhoki (a.u.) */
    double K[4], L[4], M[4];            /* eV, 1/A^2, A                */
} orc_raw;

typedef struct {
    int Z, nao, nshell;       /* nshell: principal quantum number of valence */
    double core;              /* core charge                                 */
    double eisol, dhf;        /* MNDO atomic electronic energy (eV), dHf (kcal/mol) */
    orc_raw th[3];            /* MNDO, AM1, PM3                              */
} orc_raw_elem;

static const orc_raw_elem RAW[] = {
  /* Hatom.cpp:50-220 */
  { 1, 1, 1, 1.0, -11.906276, 52.102, {
    { -11.906276, 0.0, 1.331967, 0.0, -6.989064, 0.0, 2.544134,
      12.848, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.4721515374, 0.0, 0.0,
      {0,0,0,0}, {0,0,0,0}, {0,0,0,0} },
    { -11.396427, 0.0, 1.188078, 0.0, -6.173787, 0.0, 2.882324,
      12.848, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.4721515374, 0.0, 0.0,
      {0.122796, 0.005090, -0.018336, 0.0}, {5.0, 5.0, 2.0, 0.0}, {1.2, 1.8, 2.1, 0.0} },
    { -13.073321, 0.0, 0.967807, 0.0, -5.626512, 0.0, 3.356386,
      14.794208, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.5436727936, 0.0, 0.0,
      {1.128750, -1.060329, 0.0, 0.0}, {5.096282, 6.003788, 0.0, 0.0}, {1.537465, 1.570189, 0.0, 0.0} } } },
  /* Catom.cpp:47-225 */
  { 6, 4, 2, 4.0, -120.500606, 170.89, {
    { -52.279745, -39.205558, 1.787537, 1.787537, -18.985044, -7.934122, 2.546380,
      12.23, 11.08, 11.47, 9.84, 2.43, 0.8074661800, 0.6851577737, 0.4494406369, 0.6149309919, 0.6685771472,
      {0,0,0,0}, {0,0,0,0}, {0,0,0,0} },
    { -52.028658, -39.614239, 1.808665, 1.685116, -15.715783, -7.719283, 2.648274,
      12.23, 11.08, 11.47, 9.84, 2.43, 0.8236735591, 0.7268015207, 0.4494406369, 0.6082783276, 0.6423370115,
      {0.011355, 0.045924, -0.020061, -0.001260}, {5.0, 5.0, 5.0, 5.0}, {1.60, 1.85, 2.05, 2.65} },
    { -47.270320, -36.266918, 1.565085, 1.842345, -11.910015, -9.802755, 2.707807,
      11.200708, 10.796292, 10.265027, 9.042566, 2.290980, 0.8332396384, 0.6647749859, 0.4116151543, 0.5885706542, 0.7647513703,
      {0.050107, 0.050733, 0.0, 0.0}, {6.003165, 6.002979, 0.0, 0.0}, {1.642214, 0.892488, 0.0, 0.0} } } },
  /* Natom.cpp */
  { 7, 4, 2, 5.0, -202.581201, 113.00, {
    { -71.932122, -57.172319, 2.255614, 2.255614, -20.495758, -20.495758, 2.861342,
      13.59, 12.98, 12.66, 11.59, 3.14, 0.6399036683, 0.5429762678, 0.4994193177, 0.7843433156, 0.8126295047,
      {0,0,0,0}, {0,0,0,0}, {0,0,0,0} },
    { -71.860000, -57.167581, 2.315410, 2.157940, -20.299110, -18.238666, 2.947286,
      13.59, 12.98, 12.66, 11.59, 3.14, 0.6433247425, 0.5675527917, 0.4994193177, 0.7820630445, 0.7883351388,
      {0.025251, 0.028953, -0.005806, 0.0}, {5.0, 5.0, 2.0, 0.0}, {1.50, 2.10, 2.40, 0.0} },
    { -49.335672, -47.509736, 2.028094, 2.313728, -14.062521, -20.043848, 2.830545,
      11.904787, 11.754672, 7.348565, 10.807277, 1.136713, 0.6577005762, 0.5293383109, 0.4374893746, 0.5030877737, 0.7364801616,
      {1.501674, -1.505772, 0.0, 0.0}, {5.901148, 6.004658, 0.0, 0.0}, {1.710740, 1.716149, 0.0, 0.0} } } },
  /* Oatom.cpp:47-225 */
  { 8, 4, 2, 6.0, -317.868506, 59.559, {
    { -99.64309, -77.797472, 2.699905, 2.699905, -32.688082, -32.688082, 3.160604,
      15.42, 14.52, 14.48, 12.98, 3.94, 0.5346023927, 0.4536251725, 0.5666700426, 0.9592303457, 0.9495760934,
      {0,0,0,0}, {0,0,0,0}, {0,0,0,0} },
    { -97.830000, -78.262380, 3.108032, 2.524039, -29.272773, -29.272773, 4.455371,
      15.42, 14.52, 14.48, 12.98, 3.94, 0.4988896404, 0.4852321503, 0.5666700426, 0.9960801167, 0.9065055775,
      {0.280962, 0.081430, 0.0, 0.0}, {5.0, 7.0, 0.0, 0.0}, {0.847918, 1.445071, 0.0, 0.0} },
    { -86.993002, -71.879580, 3.796544, 2.389402, -45.202651, -24.752515, 3.217102,
      15.755760, 13.654016, 10.621160, 12.40609, 0.593883, 0.4086173087, 0.5125738036, 0.5790088969, 0.5299517372, 0.8179482975,
      {-1.131128, 1.137891, 0.0, 0.0}, {6.002477, 5.950512, 0.0, 0.0}, {1.607311, 1.598395, 0.0, 0.0} } } },
  /* Fatom.cpp */
  { 9, 4, 2, 7.0, -476.683781, 18.86, {
    { -131.071548, -105.782137, 2.848487, 2.848487, -48.290460, -36.508540, 3.419661,
      16.92, 16.71, 17.25, 14.91, 4.83, 0.5067166088, 0.4299633003, 0.6217935876, 1.0850000098, 1.0343451703,
      {0,0,0,0}, {0,0,0,0}, {0,0,0,0} },
    { -136.105579, -104.889885, 3.770082, 2.494670, -69.590277, -27.922360, 5.517800,
      16.92, 16.71, 17.25, 14.91, 4.83, 0.4145203025, 0.4909446425, 0.6217935876, 1.2088469198, 0.9449175360,
      {0.242079, 0.003607, 0.0, 0.0}, {4.80, 4.60, 0.0, 0.0}, {0.930, 1.660, 0.0, 0.0} },
    { -110.435303, -105.685047, 4.708555, 2.491178, -48.405939, -27.744660, 3.358921,
      10.496667, 14.817256, 16.073689, 14.418393, 0.727763, 0.3125302275, 0.4916328225, 0.3857423305, 0.6768359077, 0.6119953427,
      {-0.012166, -0.002852, 0.0, 0.0}, {6.023574, 6.003717, 0.0, 0.0}, {1.856859, 2.636158, 0.0, 0.0} } } },
};
#define ORC_NRAW ((int)(sizeof(RAW) / sizeof(RAW[0])))

/* Converted per-element parameters (atomic units). */
typedef struct {
    int Z, nao, nshell;
    double core, U[2], zeta[2], beta[2], alpha;
    double Gss, Gpp, Gsp, Gpp2, Hsp, Hpp;
    double D[3], rho[3];
    double K[4], L[4], M[4];
    double eisol, dhf;
} orc_elem;

static int orc_elem_params(int Z, int theory, orc_elem* e)
{
    const orc_raw_elem* r = NULL;
    for (int i = 0; i < ORC_NRAW; i++)
        if (RAW[i].Z == Z) r = &RAW[i];
    if (!r || theory < 0 || theory > 2) return -1;
    const orc_raw* p = &r->th[theory];
    memset(e, 0, sizeof(*e));
    e->Z = Z; e->nao = r->nao; e->nshell = r->nshell; e->core = r->core;
    e->U[0] = p->Uss * ORC_EV2AU;   e->U[1] = p->Upp * ORC_EV2AU;
    e->zeta[0] = p->zs;             e->zeta[1] = p->zp;
    e->beta[0] = p->bs * ORC_EV2AU; e->beta[1] = p->bp * ORC_EV2AU;
    e->alpha = p->alpha / ORC_ANG2AU;
    e->Gss = p->Gss * ORC_EV2AU;  e->Gpp = p->Gpp * ORC_EV2AU;
    e->Gsp = p->Gsp * ORC_EV2AU;  e->Gpp2 = p->Gpp2 * ORC_EV2AU;
    e->Hsp = p->Hsp * ORC_EV2AU;
    e->Hpp = 0.5 * (e->Gpp - e->Gpp2);                   /* Atom.cpp:1051-1070 */
    e->D[0] = 0.0; e->D[1] = p->D1; e->D[2] = p->D2;
    e->rho[0] = 0.5 / p->rho0i;
    e->rho[1] = (p->rho1i != 0.0) ? 0.5 / p->rho1i : 0.0;
    e->rho[2] = (p->rho2i != 0.0) ? 0.5 / p->rho2i : 0.0;
    for (int k = 0; k < 4; k++) {
        e->K[k] = p->K[k] * ORC_EV2AU;
        e->L[k] = p->L[k] / pow(ORC_ANG2AU, 2.0);
        e->M[k] = p->M[k] * ORC_ANG2AU;
    }
    e->eisol = r->eisol * ORC_EV2AU;
    e->dhf = r->dhf * ORC_KCAL2AU;
    return 0;
}

/* ------------------------------------------------------------ molecule */
typedef struct {
    int theory, natoms, nao, nelec;
    orc_elem* el;      /* [natoms] */
    int* first;        /* [natoms] first AO index */
    double* xyz;       /* [natoms*3] bohr */
    double* dist;      /* [natoms*natoms] */
    double* S;         /* [nao*nao] */
    double* g;         /* dense (mu nu|lam sig)_{AB}: [N][N][4][4][4][4] */
    double ecore;
} orc_mol;

#define G6(m, A, B, i, j, k, l) \
    ((m)->g[((((((size_t)(A) * (m)->natoms + (B)) * 4 + (i)) * 4 + (j)) * 4 + (k)) * 4 + (l))])

/* ---------------------------------------------------- rotation matrix */
/* Cndo2.cpp:5498-5531 with EularAngle.cpp:38-61 (gamma = 0). R[new][old]. */
static void orc_rotation(const double* xa, const double* xb, double R[4][4])
{
    double x = xb[0] - xa[0], y = xb[1] - xa[1], z = xb[2] - xa[2];
    double r = sqrt(pow(x, 2.0) + pow(y, 2.0) + pow(z, 2.0));
    double beta = acos(z / r), alpha = 0.0;
    if (!(x == 0.0 && y == 0.0)) {
        double rxy = sqrt(pow(x, 2.0) + pow(y, 2.0));
        alpha = atan2(y / rxy, x / rxy);
    }
    memset(R, 0, sizeof(double) * 16);
    R[S][S] = 1.0;
    R[PY][PY] = cos(alpha); R[PY][PZ] = sin(alpha) * sin(beta); R[PY][PX] = sin(alpha) * cos(beta);
    R[PZ][PY] = 0.0;        R[PZ][PZ] = cos(beta);              R[PZ][PX] = -1.0 * sin(beta);
    R[PX][PY] = -sin(alpha); R[PX][PZ] = cos(alpha) * sin(beta); R[PX][PX] = cos(alpha) * cos(beta);
}

/* ------------------------------------------------------------ overlap */
/* A_k(p) = int_1^inf x^k e^{-p x} dx  (Pople-Beveridge B.22, Cndo2.cpp:6385-6399) */
static double orc_fact(int n) { double f = 1.0; for (int i = 2; i <= n; i++) f *= i; return f; }

double orc_aux_A(int k, double rho)
{
    double s = 0.0;
    for (int mu = 1; mu <= k + 1; mu++)
        s += 1.0 / (pow(rho, mu) * orc_fact(k - mu + 1));
    return exp(-rho) * orc_fact(k) * s;
}

/* B_k(q) = int_{-1}^{1} x^k e^{-q x} dx  (B.23, Cndo2.cpp:6412-6447) */
double orc_aux_B(int k, double rho)
{
    if (fabs(rho) > 0.0) {
        double t1 = 0.0, t2 = 0.0;
        for (int mu = 1; mu <= k + 1; mu++) {
            double c = orc_fact(k) / orc_fact(k - mu + 1);
            double sg = ((k - mu) % 2 == 0) ? 1.0 : -1.0;
            t1 += c / pow(rho, mu);
            t2 += c * sg / pow(rho, mu);
        }
        return -exp(-rho) * t1 - exp(rho) * t2;
    }
    return (k % 2 == 0) ? 2.0 / (1.0 + k) : 0.0;
}

/* Polynomial in (xi, eta) of degree < 8 each: c[i][j] xi^i eta^j. */
typedef struct { double c[8][8]; } orc_poly;

static void poly_one(orc_poly* p) { memset(p, 0, sizeof(*p)); p->c[0][0] = 1.0; }
static void poly_mul(orc_poly* p, const orc_poly* q)
{
    orc_poly r; memset(&r, 0, sizeof(r));
    for (int i = 0; i < 8; i++) for (int j = 0; j < 8; j++) if (p->c[i][j] != 0.0)
        for (int k = 0; i + k < 8; k++) for (int l = 0; j + l < 8; l++)
            r.c[i + k][j + l] += p->c[i][j] * q->c[k][l];
    *p = r;
}
static void poly_mul_terms(orc_poly* p, int n, const double* cx) /* cx: triples (coef, i, j) */
{
    orc_poly q; memset(&q, 0, sizeof(q));
    for (int t = 0; t < n; t++) q.c[(int)cx[3 * t + 1]][(int)cx[3 * t + 2]] += cx[3 * t];
    poly_mul(p, &q);
}

/* Integrand polynomial of the diatomic-frame overlap between an orbital of
 * principal number na, angular l_a on A (at z = -R/2) and nb, l_b on B
 * (z = +R/2), same |m|, both with z along A->B:
 *   r_a = R/2 (xi+eta), r_b = R/2 (xi-eta), z-z_A = R/2 (1+xi eta),
 *   z-z_B = R/2 (xi eta-1), rho_perp^2 = (R/2)^2 (xi^2-1)(1-eta^2),
 *   dV = (R/2)^3 (xi^2-eta^2) dxi deta dphi.
 * Returns the angular factor (spherical-harmonic norms times the phi integral). */
double orc_overlap_poly(int na, int la, int nb, int lb, int m, orc_poly* p)
{
    const double plus[] = { 1.0, 1, 0, 1.0, 0, 1 };        /* xi + eta  */
    const double minus[] = { 1.0, 1, 0, -1.0, 0, 1 };      /* xi - eta  */
    const double zA[] = { 1.0, 0, 0, 1.0, 1, 1 };         /* 1 + xi eta */
    const double zB[] = { 1.0, 1, 1, -1.0, 0, 0 };        /* xi eta - 1 */
    const double perp[] = { 1.0, 2, 0, -1.0, 2, 2, -1.0, 0, 0, 1.0, 0, 2 }; /* (xi^2-1)(1-eta^2) */
    const double vol[] = { 1.0, 2, 0, -1.0, 0, 2 };        /* xi^2 - eta^2 */
    poly_one(p);
    for (int i = 0; i < na - 1 - la; i++) poly_mul_terms(p, 2, plus);
    for (int i = 0; i < nb - 1 - lb; i++) poly_mul_terms(p, 2, minus);
    double ang = 1.0;
    const double ns = 1.0 / sqrt(4.0 * M_PI), np = sqrt(3.0 / (4.0 * M_PI));
    ang *= (la == 0) ? ns : np;
    ang *= (lb == 0) ? ns : np;
    if (m == 0) {
        if (la == 1) poly_mul_terms(p, 2, zA);
        if (lb == 1) poly_mul_terms(p, 2, zB);
        ang *= 2.0 * M_PI;
    } else {
        poly_mul_terms(p, 4, perp);
        ang *= M_PI;
    }
    poly_mul_terms(p, 2, vol);
    return ang;
}

/* Diatomic-frame overlap between orbitals (na,la,|m|,zeta_a) and (nb,lb,|m|,zeta_b)
 * at distance R (Cndo2.cpp:5801-5869). */
double orc_overlap_local(int na, int la, double za, int nb, int lb, double zb, int m, double R)
{
    orc_poly p;
    double ang = orc_overlap_poly(na, la, nb, lb, m, &p);
    double al = za * R, be = zb * R, s = 0.0;
    for (int i = 0; i < 8; i++) for (int j = 0; j < 8; j++) if (p.c[i][j] != 0.0)
        s += p.c[i][j] * orc_aux_A(i, 0.5 * (al + be)) * orc_aux_B(j, 0.5 * (al - be));
    double pre = pow(2.0 * za, na + 0.5) * pow(2.0 * zb, nb + 0.5);
    pre /= sqrt(orc_fact(2 * na) * orc_fact(2 * nb));
    pre *= pow(R / 2.0, na + nb + 1.0);
    return pre * ang * s;
}

/* -------------------------------------------------------- multipoles */
enum { MQ = 0, MUX, MUY, MUZ, QXX, QYY, QZZ, QXZ, QYZ, QXY, NMULT };
static int mp_kind(int t) { return t == MQ ? 0 : (t <= MUZ ? 1 : 2); }

/* Semi-empirical multipole-multipole interaction (Mndo.cpp:6592-6834;
 * Dewar & Thiel, Theor. Chim. Acta 46, 89 (1977), eqs. 52-66). */
double orc_multipole(const orc_elem* A, const orc_elem* B, int ma, int mb, double r)
{
    double DA = A->D[mp_kind(ma)], DB = B->D[mp_kind(mb)];
    double a = A->rho[mp_kind(ma)] + B->rho[mp_kind(mb)], a2 = a * a;
#define IR(x) (1.0 / sqrt(x))
    if (ma == MQ && mb == MQ) return IR(r * r + a2);
    if (ma == MQ && mb == MUZ)
        return IR((r + DB) * (r + DB) + a2) / 2.0 - IR((r - DB) * (r - DB) + a2) / 2.0;
    if (ma == MUZ && mb == MQ) return -orc_multipole(B, A, mb, ma, r);
    if (ma == MQ && mb == QXX)
        return IR(r * r + 4.0 * DB * DB + a2) / 2.0 - IR(r * r + a2) / 2.0;
    if ((ma == QXX || ma == QYY) && mb == MQ) return orc_multipole(B, A, mb, ma, r);
    if (ma == MQ && mb == QYY) return orc_multipole(A, B, ma, QXX, r);
    if (ma == MQ && mb == QZZ)
        return IR((r + 2.0 * DB) * (r + 2.0 * DB) + a2) / 4.0 - IR(r * r + a2) / 2.0
             + IR((r - 2.0 * DB) * (r - 2.0 * DB) + a2) / 4.0;
    if (ma == QZZ && mb == MQ) return orc_multipole(B, A, mb, ma, r);
    if (ma == MUX && mb == MUX)
        return IR(r * r + (DA - DB) * (DA - DB) + a2) / 2.0 - IR(r * r + (DA + DB) * (DA + DB) + a2) / 2.0;
    if (ma == MUY && mb == MUY) return orc_multipole(A, B, MUX, MUX, r);
    if (ma == MUZ && mb == MUZ)
        return IR((r + DA - DB) * (r + DA - DB) + a2) / 4.0 - IR((r + DA + DB) * (r + DA + DB) + a2) / 4.0
             - IR((r - DA - DB) * (r - DA - DB) + a2) / 4.0 + IR((r - DA + DB) * (r - DA + DB) + a2) / 4.0;
    if (ma == MUX && mb == QXZ)
        return -IR((r - DB) * (r - DB) + (DA - DB) * (DA - DB) + a2) / 4.0
             + IR((r - DB) * (r - DB) + (DA + DB) * (DA + DB) + a2) / 4.0
             + IR((r + DB) * (r + DB) + (DA - DB) * (DA - DB) + a2) / 4.0
             - IR((r + DB) * (r + DB) + (DA + DB) * (DA + DB) + a2) / 4.0;
    if (ma == QXZ && mb == MUX) return -orc_multipole(B, A, mb, ma, r);
    if (ma == MUY && mb == QYZ) return orc_multipole(A, B, MUX, QXZ, r);
    if (ma == QYZ && mb == MUY) return -orc_multipole(B, A, mb, ma, r);
    if (ma == MUZ && mb == QXX)
        return -IR((r + DA) * (r + DA) + 4.0 * DB * DB + a2) / 4.0
             + IR((r - DA) * (r - DA) + 4.0 * DB * DB + a2) / 4.0
             + IR((r + DA) * (r + DA) + a2) / 4.0 - IR((r - DA) * (r - DA) + a2) / 4.0;
    if ((ma == QXX || ma == QYY) && mb == MUZ) return -orc_multipole(B, A, mb, ma, r);
    if (ma == MUZ && mb == QYY) return orc_multipole(A, B, MUZ, QXX, r);
    if (ma == MUZ && mb == QZZ)
        return -IR((r + DA - 2.0 * DB) * (r + DA - 2.0 * DB) + a2) / 8.0
             + IR((r - DA - 2.0 * DB) * (r - DA - 2.0 * DB) + a2) / 8.0
             - IR((r + DA + 2.0 * DB) * (r + DA + 2.0 * DB) + a2) / 8.0
             + IR((r - DA + 2.0 * DB) * (r - DA + 2.0 * DB) + a2) / 8.0
             + IR((r + DA) * (r + DA) + a2) / 4.0 - IR((r - DA) * (r - DA) + a2) / 4.0;
    if (ma == QZZ && mb == MUZ) return -orc_multipole(B, A, mb, ma, r);
    if ((ma == QXX && mb == QXX) || (ma == QYY && mb == QYY)) {
        return IR(r * r + 4.0 * (DA - DB) * (DA - DB) + a2) / 8.0
             + IR(r * r + 4.0 * (DA + DB) * (DA + DB) + a2) / 8.0
             - IR(r * r + 4.0 * DA * DA + a2) / 4.0 - IR(r * r + 4.0 * DB * DB + a2) / 4.0
             + IR(r * r + a2) / 4.0;
    }
    if (ma == QXX && mb == QYY)
        return IR(r * r + 4.0 * DA * DA + 4.0 * DB * DB + a2) / 4.0 - IR(r * r + 4.0 * DA * DA + a2) / 4.0
             - IR(r * r + 4.0 * DB * DB + a2) / 4.0 + IR(r * r + a2) / 4.0;
    if (ma == QYY && mb == QXX) return orc_multipole(B, A, mb, ma, r);
    if (ma == QXX && mb == QZZ)
        return IR((r - 2.0 * DB) * (r - 2.0 * DB) + 4.0 * DA * DA + a2) / 8.0
             + IR((r + 2.0 * DB) * (r + 2.0 * DB) + 4.0 * DA * DA + a2) / 8.0
             - IR((r - 2.0 * DB) * (r - 2.0 * DB) + a2) / 8.0 - IR((r + 2.0 * DB) * (r + 2.0 * DB) + a2) / 8.0
             - IR(r * r + 4.0 * DA * DA + a2) / 4.0 + IR(r * r + a2) / 4.0;
    if (ma == QZZ && (mb == QXX || mb == QYY)) return orc_multipole(B, A, mb, ma, r);
    if (ma == QYY && mb == QZZ) return orc_multipole(A, B, QXX, mb, r);
    if (ma == QZZ && mb == QZZ)
        return IR((r + 2.0 * DA - 2.0 * DB) * (r + 2.0 * DA - 2.0 * DB) + a2) / 16.0
             + IR((r + 2.0 * DA + 2.0 * DB) * (r + 2.0 * DA + 2.0 * DB) + a2) / 16.0
             + IR((r - 2.0 * DA - 2.0 * DB) * (r - 2.0 * DA - 2.0 * DB) + a2) / 16.0
             + IR((r - 2.0 * DA + 2.0 * DB) * (r - 2.0 * DA + 2.0 * DB) + a2) / 16.0
             - IR((r + 2.0 * DA) * (r + 2.0 * DA) + a2) / 8.0 - IR((r - 2.0 * DA) * (r - 2.0 * DA) + a2) / 8.0
             - IR((r + 2.0 * DB) * (r + 2.0 * DB) + a2) / 8.0 - IR((r - 2.0 * DB) * (r - 2.0 * DB) + a2) / 8.0
             + IR(r * r + a2) / 4.0;
    if ((ma == QXZ && mb == QXZ) || (ma == QYZ && mb == QYZ)) {
        double p1 = r + DA - DB, p2 = r + DA + DB, p3 = r - DA - DB, p4 = r - DA + DB;
        double m2 = (DA - DB) * (DA - DB), q2 = (DA + DB) * (DA + DB);
        return IR(p1 * p1 + m2 + a2) / 8.0 - IR(p1 * p1 + q2 + a2) / 8.0
             - IR(p2 * p2 + m2 + a2) / 8.0 + IR(p2 * p2 + q2 + a2) / 8.0
             - IR(p3 * p3 + m2 + a2) / 8.0 + IR(p3 * p3 + q2 + a2) / 8.0
             + IR(p4 * p4 + m2 + a2) / 8.0 - IR(p4 * p4 + q2 + a2) / 8.0;
    }
    if (ma == QXY && mb == QXY)
        return IR(r * r + 2.0 * (DA - DB) * (DA - DB) + a2) / 4.0
             + IR(r * r + 2.0 * (DA + DB) * (DA + DB) + a2) / 4.0
             - IR(r * r + 2.0 * DA * DA + 2.0 * DB * DB + a2) / 2.0;
#undef IR
    return 0.0; /* symmetry-forbidden combination */
}

/* Charge distribution (mu nu) on one atom as a sum of multipoles (local frame). */
static int orc_dist(int mu, int nu, int* out)
{
    if (mu > nu) { int t = mu; mu = nu; nu = t; }
    if (mu == S && nu == S) { out[0] = MQ; return 1; }
    if (mu == S && nu == PZ) { out[0] = MUZ; return 1; }
    if (mu == S && nu == PX) { out[0] = MUX; return 1; }
    if (mu == S && nu == PY) { out[0] = MUY; return 1; }
    if (mu == PZ && nu == PZ) { out[0] = MQ; out[1] = QZZ; return 2; }
    if (mu == PX && nu == PX) { out[0] = MQ; out[1] = QXX; return 2; }
    if (mu == PY && nu == PY) { out[0] = MQ; out[1] = QYY; return 2; }
    if (mu == PZ && nu == PX) { out[0] = QXZ; return 1; }
    if (mu == PY && nu == PZ) { out[0] = QYZ; return 1; }
    out[0] = QXY; return 1;     /* (py px) */
}

/* Diatomic-frame (mu nu|lam sig), Mndo.cpp:4886-5238 (eqs. 29-49). */
double orc_local_integral(const orc_elem* A, int mu, int nu, const orc_elem* B, int lam, int sig, double r)
{
    int isxy_a = ((mu == PX && nu == PY) || (mu == PY && nu == PX));
    int isxy_b = ((lam == PX && sig == PY) || (lam == PY && sig == PX));
    if (isxy_a && isxy_b)   /* eq. 49: (px py|px py) = 1/2[(px px|px px) - (px px|py py)] */
        return 0.5 * (orc_local_integral(A, PX, PX, B, PX, PX, r) - orc_local_integral(A, PX, PX, B, PY, PY, r));
    if (isxy_a || isxy_b) return 0.0;
    int da[2], db[2];
    int na = orc_dist(mu, nu, da), nb = orc_dist(lam, sig, db);
    double v = 0.0;
    for (int i = 0; i < na; i++)
        for (int j = 0; j < nb; j++) v += orc_multipole(A, B, da[i], db[j], r);
    return v;
}

/* ------------------------------------------------------------ setup */
static void orc_free(orc_mol* m)
{
    if (!m) return;
    free(m->el); free(m->first); free(m->xyz); free(m->dist); free(m->S); free(m->g);
    free(m);
}

void orc_destroy(void* h) { orc_free((orc_mol*)h); }

/* Builds everything DoSCF computes once per geometry (Cndo2.cpp:612-615):
 * overlap S, the dense two-centre tensor and the core repulsion energy.
 * Returns NULL with *err set: 1 unsupported element, 2 odd electrons. */
void* orc_create(int theory, int natoms, const int* Z, const double* xyz_bohr, int* err)
{
    *err = 0;
    orc_mol* m = (orc_mol*)calloc(1, sizeof(orc_mol));
    m->theory = theory; m->natoms = natoms;
    m->el = (orc_elem*)calloc(natoms, sizeof(orc_elem));
    m->first = (int*)calloc(natoms, sizeof(int));
    m->xyz = (double*)malloc(sizeof(double) * 3 * natoms);
    memcpy(m->xyz, xyz_bohr, sizeof(double) * 3 * natoms);
    int nao = 0, nel = 0;
    for (int a = 0; a < natoms; a++) {
        if (orc_elem_params(Z[a], theory, &m->el[a]) != 0) { *err = 1; orc_free(m); return NULL; }
        m->first[a] = nao;
        nao += m->el[a].nao;
        nel += (int)m->el[a].core;
    }
    if (nel % 2 != 0) { *err = 2; orc_free(m); return NULL; }
    m->nao = nao; m->nelec = nel;
    m->dist = (double*)calloc((size_t)natoms * natoms, sizeof(double));
    for (int a = 0; a < natoms; a++)
        for (int b = a; b < natoms; b++) {
            const double* pa = &xyz_bohr[3 * a]; const double* pb = &xyz_bohr[3 * b];
            double d = sqrt(pow(pa[0] - pb[0], 2.0) + pow(pa[1] - pb[1], 2.0) + pow(pa[2] - pb[2], 2.0));
            m->dist[a * natoms + b] = d; m->dist[b * natoms + a] = d;
        }

    /* overlap, Cndo2.cpp:4030-4146; atom pairs in parallel (each writes its own block) */
    m->S = (double*)calloc((size_t)nao * nao, sizeof(double));
#pragma omp parallel for schedule(dynamic)
    for (int a = 0; a < natoms; a++)
        for (int b = a + 1; b < natoms; b++) {
            const orc_elem *A = &m->el[a], *B = &m->el[b];
            double r = m->dist[a * natoms + b], loc[4][4], R[4][4], rot[4][4];
            memset(loc, 0, sizeof(loc));
            /* m-matched pairs in the diatomic frame: (s|s),(s|pz),(pz|s),(pz|pz),(px|px),(py|py) */
            for (int i = 0; i < A->nao; i++)
                for (int j = 0; j < B->nao; j++) {
                    int la = (i == S) ? 0 : 1, lb = (j == S) ? 0 : 1;
                    int ma = (i == PX) ? 1 : (i == PY ? -1 : 0), mb = (j == PX) ? 1 : (j == PY ? -1 : 0);
                    if (ma != mb) continue;
                    loc[i][j] = orc_overlap_local(A->nshell, la, A->zeta[la], B->nshell, lb, B->zeta[lb], abs(ma), r);
                }
            orc_rotation(&m->xyz[3 * a], &m->xyz[3 * b], R);
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    double s = 0.0;
                    for (int k = 0; k < 4; k++)
                        for (int l = 0; l < 4; l++) s += R[i][k] * loc[k][l] * R[j][l];
                    rot[i][j] = s;
                }
            for (int i = 0; i < A->nao; i++)
                for (int j = 0; j < B->nao; j++) {
                    int mu = m->first[a] + i, nu = m->first[b] + j;
                    m->S[(size_t)mu * nao + nu] = rot[i][j];
                    m->S[(size_t)nu * nao + mu] = rot[i][j];
                }
        }
    for (int mu = 0; mu < nao; mu++) m->S[(size_t)mu * nao + mu] = 1.0;

    /* two-centre integrals, Mndo.cpp:3703-3820, 3948-4025, 4223-4293 */
    m->g = (double*)calloc((size_t)natoms * natoms * 256, sizeof(double));
#pragma omp parallel for schedule(dynamic)
    for (int a = 0; a < natoms; a++)
        for (int b = a + 1; b < natoms; b++) {
            const orc_elem *A = &m->el[a], *B = &m->el[b];
            double r = m->dist[a * natoms + b], R[4][4];
            double loc[16][16], T[16][16], tmp[16][16], out[16][16];
            memset(loc, 0, sizeof(loc));
            for (int mu = 0; mu < A->nao; mu++)
                for (int nu = mu; nu < A->nao; nu++)
                    for (int la = 0; la < B->nao; la++)
                        for (int sg = la; sg < B->nao; sg++) {
                            double v = orc_local_integral(A, mu, nu, B, la, sg, r);
                            loc[mu * 4 + nu][la * 4 + sg] = v; loc[mu * 4 + nu][sg * 4 + la] = v;
                            loc[nu * 4 + mu][la * 4 + sg] = v; loc[nu * 4 + mu][sg * 4 + la] = v;
                        }
            orc_rotation(&m->xyz[3 * a], &m->xyz[3 * b], R);
            for (int i = 0; i < 16; i++)
                for (int j = 0; j < 16; j++) T[i][j] = R[i / 4][j / 4] * R[i % 4][j % 4];
            for (int i = 0; i < 16; i++)
                for (int j = 0; j < 16; j++) {
                    double s = 0.0;
                    for (int k = 0; k < 16; k++) s += T[i][k] * loc[k][j];
                    tmp[i][j] = s;
                }
            for (int i = 0; i < 16; i++)
                for (int j = 0; j < 16; j++) {
                    double s = 0.0;
                    for (int k = 0; k < 16; k++) s += tmp[i][k] * T[j][k];
                    out[i][j] = s;
                }
            /* 8-fold fill of [a][b] and [b][a] from the mu<=nu, lam<=sig values */
            for (int mu = 0; mu < 4; mu++)
                for (int nu = mu; nu < 4; nu++)
                    for (int la = 0; la < 4; la++)
                        for (int sg = la; sg < 4; sg++) {
                            double v = out[mu * 4 + nu][la * 4 + sg];
                            G6(m, a, b, mu, nu, la, sg) = v; G6(m, a, b, mu, nu, sg, la) = v;
                            G6(m, a, b, nu, mu, la, sg) = v; G6(m, a, b, nu, mu, sg, la) = v;
                            G6(m, b, a, la, sg, mu, nu) = v; G6(m, b, a, la, sg, nu, mu) = v;
                            G6(m, b, a, sg, la, mu, nu) = v; G6(m, b, a, sg, la, nu, mu) = v;
                        }
        }

    /* core repulsion, Cndo2.cpp:406-416 -> Am1.cpp:145-170 -> Mndo.cpp:223-240, 319-327 */
    double e = 0.0, ec = 0.0;
    for (int a = 0; a < natoms; a++)
        for (int b = a + 1; b < natoms; b++) {
            const orc_elem *A = &m->el[a], *B = &m->el[b];
            double r = m->dist[a * natoms + b];
            double f;
            if (A->Z == 1 && (B->Z == 7 || B->Z == 8))
                f = 1.0 + (r / ORC_ANG2AU) * exp(-B->alpha * r) + exp(-A->alpha * r);
            else if (B->Z == 1 && (A->Z == 7 || A->Z == 8))
                f = 1.0 + (r / ORC_ANG2AU) * exp(-A->alpha * r) + exp(-B->alpha * r);
            else
                f = 1.0 + exp(-A->alpha * r) + exp(-B->alpha * r);
            double term = A->core * B->core * G6(m, a, b, S, S, S, S) * f;
            if (theory != ORC_MNDO) {
                double t = 0.0;
                for (int k = 0; k < 4; k++) {
                    t += A->K[k] * exp(-A->L[k] * pow(r - A->M[k], 2.0));
                    t += B->K[k] * exp(-B->L[k] * pow(r - B->M[k], 2.0));
                }
                term += A->core * B->core * t * ORC_ANG2AU / r;
            }
            nsum_add(&e, &ec, term);
        }
    m->ecore = e + ec;
    return m;
}

int orc_nao(void* h) { return ((orc_mol*)h)->nao; }
int orc_nelec(void* h) { return ((orc_mol*)h)->nelec; }
double orc_ecore(void* h) { return ((orc_mol*)h)->ecore; }
void orc_get_overlap(void* h, double* out)
{
    orc_mol* m = (orc_mol*)h;
    memcpy(out, m->S, sizeof(double) * m->nao * m->nao);
}
/* Dense tensor copy: out[N][N][4][4][4][4]. */
void orc_get_integrals(void* h, double* out)
{
    orc_mol* m = (orc_mol*)h;
    memcpy(out, m->g, sizeof(double) * (size_t)m->natoms * m->natoms * 256);
}

/* One-centre Coulomb g and exchange h (Mndo.cpp:583-639). */
static double orc_g1(const orc_elem* e, int a, int b)
{
    if (a == S && b == S) return e->Gss;
    if (a == S || b == S) return e->Gsp;
    return (a == b) ? e->Gpp : e->Gpp2;
}
static double orc_h1(const orc_elem* e, int a, int b)
{
    if (a == b) return orc_g1(e, a, b);
    if (a == S || b == S) return e->Hsp;
    return e->Hpp;
}

/* Fock matrix (Cndo2.cpp:1526-1636 with Mndo.cpp:422-650).  The upper triangle
 * is computed element by element as the reference does, then mirrored; the
 * atom rows A in parallel as the reference's OpenMP loop (Cndo2.cpp:1555),
 * each element by one thread in a fixed order, so the result does not depend
 * on the thread count.  P is n x n row-major, F is written as a full
 * symmetric n x n. */
void orc_fock(void* h, const double* P, int is_guess, double* F)
{
    orc_mol* m = (orc_mol*)h;
    int n = m->nao, N = m->natoms;
    memset(F, 0, sizeof(double) * n * n);
#pragma omp parallel for schedule(dynamic)
    for (int A = 0; A < N; A++) {
        const orc_elem* eA = &m->el[A];
        int fA = m->first[A];
        for (int i = 0; i < eA->nao; i++) {
            int mu = fA + i;
            for (int B = A; B < N; B++) {
                const orc_elem* eB = &m->el[B];
                int fB = m->first[B];
                for (int j = 0; j < eB->nao; j++) {
                    int nu = fB + j;
                    if (nu < mu) continue;
                    double v = 0.0;
                    if (mu == nu) {
                        v = eA->U[i == S ? 0 : 1];
                        if (!is_guess) {
                            double t = 0.0;
                            for (int k = 0; k < eA->nao; k++)
                                t += P[(size_t)(fA + k) * n + fA + k] * (orc_g1(eA, i, k) - 0.5 * orc_h1(eA, i, k));
                            v += t;
                            t = 0.0;
                            for (int C = 0; C < N; C++) {
                                if (C == A) continue;
                                const orc_elem* eC = &m->el[C];
                                int fC = m->first[C];
                                for (int l = 0; l < eC->nao; l++)
                                    for (int s = 0; s < eC->nao; s++)
                                        t += P[(size_t)(fC + l) * n + fC + s] * G6(m, A, C, i, i, l, s);
                                t += -1.0 * eC->core * G6(m, A, C, i, i, S, S);
                            }
                            v += t;
                        }
                    } else {
                        double bp = 0.5 * (eA->beta[i == S ? 0 : 1] + eB->beta[j == S ? 0 : 1]);
                        if (is_guess) {
                            v = bp * m->S[(size_t)mu * n + nu];
                        } else if (A == B) {
                            double t = (1.5 * orc_h1(eA, i, j) - 0.5 * orc_g1(eA, i, j)) * P[(size_t)mu * n + nu];
                            for (int C = 0; C < N; C++) {
                                if (C == A) continue;
                                const orc_elem* eC = &m->el[C];
                                int fC = m->first[C];
                                for (int l = 0; l < eC->nao; l++)
                                    for (int s = 0; s < eC->nao; s++)
                                        t += P[(size_t)(fC + l) * n + fC + s] * G6(m, A, C, i, j, l, s);
                                t += -1.0 * eC->core * G6(m, A, C, i, j, S, S);
                            }
                            v = t;
                        } else {
                            double t = bp * m->S[(size_t)mu * n + nu];
                            for (int s = 0; s < eA->nao; s++)
                                for (int l = 0; l < eB->nao; l++)
                                    t -= 0.5 * P[(size_t)(fA + s) * n + fB + l] * G6(m, A, B, i, s, j, l);
                            v = t;
                        }
                    }
                    F[(size_t)mu * n + nu] = v;
                }
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) F[(size_t)j * n + i] = F[(size_t)i * n + j];
}

/* E_el = 1/2 sum P (F + H), F from P and H = F(P = 0), both non-guess
 * (Cndo2.cpp:1324-1416).  Returns E_el (without core repulsion). */
double orc_elec_energy(void* h, const double* P)
{
    orc_mol* m = (orc_mol*)h;
    int n = m->nao;
    double* F = (double*)malloc(sizeof(double) * n * n);
    double* H = (double*)malloc(sizeof(double) * n * n);
    double* Z = (double*)calloc((size_t)n * n, sizeof(double));
    orc_fock(h, P, 0, F);
    orc_fock(h, Z, 0, H);
    double e = 0.0, ec = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) nsum_add(&e, &ec, P[(size_t)j * n + i] * (F[(size_t)i * n + j] + H[(size_t)i * n + j]));
    free(F); free(H); free(Z);
    return 0.5 * (e + ec);
}

/* Isolated-atom electronic energy (eV) of the ground configuration s^ns p^np
 * from the one-centre parameters (Dewar & Thiel 1977; MOPAC's EISOL):
 * ns Uss + np Upp + [ns = 2] Gss + ns np Gsp - ns np Hsp / 2 + a Gpp + b Gp2,
 * (a, b) per np = 0..5: (0,0) (0,0) (-1/2,3/2) (-3/2,9/2) (-1/2,13/2) (0,10). */
double orc_eisol_ev(int Z, int theory)
{
    const orc_raw_elem* r = NULL;
    for (int i = 0; i < ORC_NRAW; i++)
        if (RAW[i].Z == Z) r = &RAW[i];
    if (!r || theory < 0 || theory > 2) return NAN;
    const orc_raw* p = &r->th[theory];
    const int ns = (Z == 1) ? 1 : 2, np = (Z == 1) ? 0 : (int)r->core - 2;
    static const double a[6] = {0.0, 0.0, -0.5, -1.5, -0.5, 0.0}, b[6] = {0.0, 0.0, 1.5, 4.5, 6.5, 10.0};
    double e = ns * p->Uss + np * p->Upp + ns * np * p->Gsp - 0.5 * ns * np * p->Hsp + a[np] * p->Gpp + b[np] * p->Gpp2;
    if (ns == 2) e += p->Gss;
    return e;
}

/* Heat of formation minus E_total: sum_A (dHf_A - E_isol_A); MNDO with the
 * tabulated E_isol (Mndo.cpp:395-404), AM1 / PM3 with orc_eisol_ev. */
double orc_hof_offset(void* h)
{
    orc_mol* m = (orc_mol*)h;
    double s = 0.0;
    for (int a = 0; a < m->natoms; a++) {
        const double eis = m->theory == 0 ? m->el[a].eisol : orc_eisol_ev(m->el[a].Z, m->theory) * ORC_EV2AU;
        s += -eis + m->el[a].dhf;
    }
    return s;
}

/* Parameter probe for tests: returns D[k], rho[k] and zeta of element Z. */
int orc_param(int Z, int theory, double* out /* D0 D1 D2 rho0 rho1 rho2 zs zp Gss Gpp Gsp Gpp2 Hsp Hpp */)
{
    orc_elem e;
    if (orc_elem_params(Z, theory, &e) != 0) return -1;
    double v[14] = { e.D[0], e.D[1], e.D[2], e.rho[0], e.rho[1], e.rho[2], e.zeta[0], e.zeta[1],
                     e.Gss, e.Gpp, e.Gsp, e.Gpp2, e.Hsp, e.Hpp };
    memcpy(out, v, sizeof(v));
    return 0;
}

/* Every converted parameter of element Z (test hook): core, U[2], zeta[2],
 * beta[2], alpha, Gss, Gpp, Gsp, Gpp2, Hsp, Hpp, D[3], rho[3], K[4], L[4],
 * M[4], and the MNDO atomic term dHf - E_isol (Mndo.cpp:395-404): 33 doubles. */
int orc_elem_dump(int Z, int theory, double* out)
{
    orc_elem e;
    if (orc_elem_params(Z, theory, &e) != 0) return -1;
    double v[33] = { e.core, e.U[0], e.U[1], e.zeta[0], e.zeta[1], e.beta[0], e.beta[1], e.alpha,
                     e.Gss, e.Gpp, e.Gsp, e.Gpp2, e.Hsp, e.Hpp, e.D[0], e.D[1], e.D[2], e.rho[0], e.rho[1], e.rho[2],
                     e.K[0], e.K[1], e.K[2], e.K[3], e.L[0], e.L[1], e.L[2], e.L[3], e.M[0], e.M[1], e.M[2], e.M[3],
                     -e.eisol + e.dhf };
    memcpy(out, v, sizeof(v));
    return 0;
}

/* Multipole probe for tests (brute-force point-charge check). */
double orc_multipole_probe(int Za, int Zb, int theory, int ma, int mb, double r)
{
    orc_elem A, B;
    if (orc_elem_params(Za, theory, &A) || orc_elem_params(Zb, theory, &B)) return NAN;
    return orc_multipole(&A, &B, ma, mb, r);
}

double orc_local_integral_probe(int Za, int Zb, int theory, int mu, int nu, int lam, int sig, double r)
{
    orc_elem A, B;
    if (orc_elem_params(Za, theory, &A) || orc_elem_params(Zb, theory, &B)) return NAN;
    return orc_local_integral(&A, mu, nu, &B, lam, sig, r);
}
