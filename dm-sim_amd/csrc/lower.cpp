// lower.cpp -- API gates -> primitive sweeps (host side).
//
// Coefficients are computed here with glibc sincos() in the same
// expressions the reference evaluates per gate (src/dmsim_amdgpu_omp.hpp
// U2/U3 :1493-1530, RX/RY :1419-1463, U1 :1467-1480), and composites expand
// into the same primitive order (:1534-1815, tool/qelib1.inc), so a plain run
// performs the reference's arithmetic sweep for sweep.
#include "lower.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <sstream>

namespace dmx {

namespace {

constexpr double kPi = 3.14159265358979323846;   // src/config.hpp:55
constexpr double kS2I = 0.70710678118654752440;  // src/config.hpp:57

// The reference's gate code takes cos and sin of the same argument in every
// coefficient set; g++ -O3 turns each pair into one glibc sincos() call (the
// compiled reference has sincos at all ten sites), and sincos can differ from
// cos() in the last bit.  Calling it explicitly gives the same doubles
// whichever compiler builds this file.
struct SinCos {
    double s, c;
    explicit SinCos(double x) { ::sincos(x, &s, &c); }
};

struct Emit {
    std::vector<PrimOp>& out;

    PrimOp& push(int kind, int q0, int q1 = -1)
    {
        PrimOp p{};
        p.kind = kind;
        p.q0 = q0;
        p.q1 = q1;
        out.push_back(p);
        return out.back();
    }
    void c1(int q, const double e[8])
    {
        PrimOp& p = push(PK_C1, q);
        for (int i = 0; i < 8; i++) p.c[i] = e[i];
    }
    void x(int q) { push(PK_X, q); }
    void y(int q) { push(PK_Y, q); }
    void z(int q) { push(PK_Z, q); }
    void h(int q) { push(PK_H, q); }
    void s(int q) { push(PK_S, q); }
    void sdg(int q) { push(PK_SDG, q); }
    void t(int q) { push(PK_T, q); }
    void tdg(int q) { push(PK_TDG, q); }
    void srn(int q) { push(PK_SRN, q); }
    void w(int q) { push(PK_W, q); }
    void r(int q, double phase) { push(PK_R, q).c[0] = phase; }
    void cx(int ctrl, int tgt) { push(PK_X, tgt, ctrl); }
    void rx(int q, double theta)
    {
        PrimOp& p = push(PK_RX, q);
        const SinCos h(theta / 2.0);
        p.c[0] = h.c;
        p.c[1] = -h.s;
    }
    void ry(int q, double theta)
    {
        PrimOp& p = push(PK_RY, q);
        const SinCos h(theta / 2.0);
        p.c[0] = h.c;
        p.c[1] = -h.s;
        p.c[2] = h.s;
        p.c[3] = h.c;
    }
    void u1(int q, double lambda)
    {
        PrimOp& p = push(PK_U1, q);
        const SinCos l(lambda);
        p.c[0] = l.c;
        p.c[1] = l.s;
    }
    void u2(int q, double phi, double lambda)
    {
        const SinCos l(lambda), p(phi), pl(phi + lambda);
        double e[8];
        e[0] = kS2I;
        e[1] = 0;
        e[2] = -kS2I * l.c;
        e[3] = -kS2I * l.s;
        e[4] = kS2I * p.c;
        e[5] = kS2I * p.s;
        e[6] = kS2I * pl.c;
        e[7] = kS2I * pl.s;
        c1(q, e);
    }
    void u3(int q, double theta, double phi, double lambda)
    {
        const SinCos h(theta / 2.), l(lambda), p(phi), pl(phi + lambda);
        double e[8];
        e[0] = h.c;
        e[1] = 0;
        e[2] = -l.c * h.s;
        e[3] = -l.s * h.s;
        e[4] = p.c * h.s;
        e[5] = p.s * h.s;
        e[6] = pl.c * h.c;
        e[7] = pl.s * h.c;
        c1(q, e);
    }
    void cu1(double lambda, int a, int b)
    {
        u1(a, lambda / 2);
        cx(a, b);
        u1(b, -lambda / 2);
        cx(a, b);
        u1(b, lambda / 2);
    }
    void ccx(int a, int b, int c)
    {
        h(c); cx(b, c); tdg(c); cx(a, c); t(c); cx(b, c); tdg(c); cx(a, c);
        t(b); t(c); h(c); cx(a, b); t(a); tdg(b); cx(a, b);
    }
    // C3X (ang = pi/4) and C3SQRTX (ang = pi/8) share one pattern.
    void c3x_like(double ang, int a, int b, int c, int d)
    {
        const int ctl[7] = {a, b, b, c, c, c, c};
        const double sgn[7] = {-1, 1, -1, 1, -1, 1, -1};
        const int cxa[6] = {a, a, b, a, b, a};
        const int cxb[6] = {b, b, c, c, c, c};
        for (int i = 0; i < 7; i++) {
            h(d);
            cu1(sgn[i] * ang, ctl[i], d);
            h(d);
            if (i < 6) cx(cxa[i], cxb[i]);
        }
    }
};

}  // namespace

const char* const kOpNames[OP_COUNT] = {
    "U3", "U2", "U1", "CX", "ID", "X", "Y", "Z", "H", "S",
    "SDG", "T", "TDG", "RX", "RY", "RZ", "CZ", "CY", "SWAP", "CH",
    "CCX", "CSWAP", "CRX", "CRY", "CRZ", "CU1", "CU3", "RXX", "RZZ", "RCCX",
    "RC3X", "C3X", "C3SQRTX", "C4X", "R", "SRN", "W", "RYY", "MAT", "KRAUS"};

int op_arity(int op)
{
    switch (op) {
    case OP_CX: case OP_CZ: case OP_CY: case OP_SWAP: case OP_CH: case OP_CRX: case OP_CRY:
    case OP_CRZ: case OP_CU1: case OP_CU3: case OP_RXX: case OP_RZZ: case OP_RYY:
        return 2;
    case OP_CCX: case OP_CSWAP: case OP_RCCX:
        return 3;
    case OP_RC3X: case OP_C3X: case OP_C3SQRTX:
        return 4;
    case OP_C4X:
        return 5;
    default:
        return 1;
    }
}

std::string Gate::dump() const
{
    std::ostringstream ss;
    ss << kOpNames[op] << "(" << qb[0] << "," << qb[1] << "," << qb[2] << "," << qb[3] << ","
       << qb[4] << "," << theta << "," << phi << "," << lambda << ");";
    return ss.str();
}

bool lower_gate(const Gate& g, std::vector<PrimOp>& out)
{
    Emit E{out};
    const int a = (int)g.qb[0], b = (int)g.qb[1], c = (int)g.qb[2], d = (int)g.qb[3],
              e = (int)g.qb[4];
    const double theta = g.theta, phi = g.phi, lambda = g.lambda;
    switch (g.op) {
    case OP_U3: E.u3(a, theta, phi, lambda); break;
    case OP_U2: E.u2(a, phi, lambda); break;
    case OP_U1: E.u1(a, lambda); break;
    case OP_CX: E.cx(a, b); break;
    case OP_ID: break;
    case OP_X: E.x(a); break;
    case OP_Y: E.y(a); break;
    case OP_Z: E.z(a); break;
    case OP_H: E.h(a); break;
    case OP_S: E.s(a); break;
    case OP_SDG: E.sdg(a); break;
    case OP_T: E.t(a); break;
    case OP_TDG: E.tdg(a); break;
    case OP_RX: E.rx(a, theta); break;
    case OP_RY: E.ry(a, theta); break;
    case OP_RZ: E.u1(a, phi); break;  // RZ keeps its angle in phi and runs U1
    case OP_CZ: E.h(b); E.cx(a, b); E.h(b); break;
    case OP_CY: E.sdg(b); E.cx(a, b); E.s(b); break;
    case OP_SWAP: E.cx(a, b); E.cx(b, a); E.cx(a, b); break;
    case OP_CH:
        E.h(b); E.sdg(b); E.cx(a, b); E.h(b); E.t(b); E.cx(a, b); E.t(b); E.h(b); E.s(b);
        E.x(b); E.s(a);
        break;
    case OP_CCX: E.ccx(a, b, c); break;
    case OP_CSWAP: E.cx(c, b); E.ccx(a, b, c); E.cx(c, b); break;
    case OP_CRX:
        E.u1(b, kPi / 2); E.cx(a, b); E.u3(b, -lambda / 2, 0, 0); E.cx(a, b);
        E.u3(b, lambda / 2, -kPi / 2, 0);
        break;
    case OP_CRY:
        E.u3(b, lambda / 2, 0, 0); E.cx(a, b); E.u3(b, -lambda / 2, 0, 0); E.cx(a, b);
        break;
    case OP_CRZ:
        E.u1(b, lambda / 2); E.cx(a, b); E.u1(b, -lambda / 2); E.cx(a, b);
        break;
    case OP_CU1: E.cu1(lambda, a, b); break;
    case OP_CU3: {
        const double t1 = (lambda - phi) / 2;
        const double t2 = theta / 2;
        const double t3 = -(phi + lambda) / 2;
        E.u1(a, -t3); E.u1(b, t1); E.cx(a, b); E.u3(b, -t2, 0, t3); E.cx(a, b);
        E.u3(b, t2, phi, 0);
        break;
    }
    case OP_RXX:
        E.u3(a, kPi / 2, theta, 0); E.h(b); E.cx(a, b); E.u1(b, -theta); E.cx(a, b); E.h(b);
        E.u2(a, -kPi, kPi - theta);
        break;
    case OP_RZZ: E.cx(a, b); E.u1(b, theta); E.cx(a, b); break;
    case OP_RCCX:
        E.u2(c, 0, kPi); E.u1(c, kPi / 4); E.cx(b, c); E.u1(c, -kPi / 4); E.cx(a, c);
        E.u1(c, kPi / 4); E.cx(b, c); E.u1(c, -kPi / 4); E.u2(c, 0, kPi);
        break;
    case OP_RC3X:
        E.u2(d, 0, kPi); E.u1(d, kPi / 4); E.cx(c, d); E.u1(d, -kPi / 4); E.u2(d, 0, kPi);
        E.cx(a, d); E.u1(d, kPi / 4); E.cx(b, d); E.u1(d, -kPi / 4); E.cx(a, d);
        E.u1(d, kPi / 4); E.cx(b, d); E.u1(d, -kPi / 4); E.u2(d, 0, kPi); E.u1(d, kPi / 4);
        E.cx(c, d); E.u1(d, -kPi / 4); E.u2(d, 0, kPi);
        break;
    case OP_C3X: E.c3x_like(kPi / 4, a, b, c, d); break;
    case OP_C3SQRTX: E.c3x_like(kPi / 8, a, b, c, d); break;
    case OP_C4X:  // qelib1 c4x as the reference spells it, including `h d; cu1 d,e; h d`
        E.h(e); E.cu1(-kPi / 2, d, e); E.h(e);
        E.c3x_like(kPi / 4, a, b, c, d);
        E.h(d); E.cu1(kPi / 4, d, e); E.h(d);
        E.c3x_like(kPi / 4, a, b, c, d);
        E.c3x_like(kPi / 8, a, b, c, e);
        break;
    case OP_R: E.r(a, theta); break;
    case OP_SRN: E.srn(a); break;
    case OP_W: E.w(a); break;
    case OP_RYY:
        E.rx(a, kPi / 2); E.rx(b, kPi / 2); E.cx(a, b); E.u1(b, theta); E.cx(a, b);
        E.rx(a, -kPi / 2); E.rx(b, -kPi / 2);
        break;
    case OP_KRAUS:
        // two records: the superoperator for each orientation of the stored rho
        for (int t = 0; t < 2; t++) {
            PrimOp& p = E.push(PK_SUP, a, a);
            p.pad = t;
            for (int i = 0; i < 16; i++) {
                p.c[2 * i] = g.mre[16 * t + i];
                p.c[2 * i + 1] = g.mim[16 * t + i];
            }
        }
        break;
    case OP_MAT:
        if (g.k == 1) {
            double e8[8];
            for (int i = 0; i < 4; i++) {
                e8[2 * i] = g.mre[i];
                e8[2 * i + 1] = g.mim[i];
            }
            E.c1(a, e8);
        } else if (g.k == 2) {
            PrimOp& p = E.push(PK_C2, a, b);
            for (int i = 0; i < 16; i++) {
                p.c[2 * i] = g.mre[i];
                p.c[2 * i + 1] = g.mim[i];
            }
        } else {
            return false;
        }
        break;
    default:
        return false;
    }
    return true;
}

namespace {

using cd = std::complex<double>;

// 2x2 complex matrix a primitive applies to (el0, el1); false for SRN, whose
// update is not complex-linear (it conjugates one operand).
bool prim_matrix2(const PrimOp& p, cd m[4])
{
    const double* c = p.c;
    const cd I(0, 1);
    switch (p.kind) {
    case PK_C1: m[0] = cd(c[0], c[1]); m[1] = cd(c[2], c[3]); m[2] = cd(c[4], c[5]);
        m[3] = cd(c[6], c[7]); return true;
    case PK_X: m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; return true;
    case PK_Y: m[0] = 0; m[1] = -I; m[2] = I; m[3] = 0; return true;
    case PK_H: m[0] = kS2I; m[1] = kS2I; m[2] = kS2I; m[3] = -kS2I; return true;
    case PK_RX: m[0] = c[0]; m[1] = I * c[1]; m[2] = I * c[1]; m[3] = c[0]; return true;
    case PK_RY: m[0] = c[0]; m[1] = c[1]; m[2] = c[2]; m[3] = c[3]; return true;
    case PK_W: m[0] = kS2I; m[1] = -I * kS2I; m[2] = -I * kS2I; m[3] = kS2I; return true;
    case PK_Z: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = -1; return true;
    case PK_S: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = I; return true;
    case PK_SDG: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = -I; return true;
    case PK_T: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = cd(kS2I, kS2I); return true;
    case PK_TDG: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = cd(kS2I, -kS2I); return true;
    case PK_R: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = I * c[0]; return true;
    case PK_U1: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = cd(c[0], c[1]); return true;
    default: return false;
    }
}

}  // namespace

void gate_matrix(const Gate& g, int& k, std::vector<uint32_t>& qubits, std::vector<double>& re,
                 std::vector<double>& im)
{
    if (g.op == OP_MAT) {
        k = g.k;
        qubits.assign(g.qb, g.qb + k);
        re = g.mre;
        im = g.mim;
        return;
    }
    k = op_arity(g.op);
    qubits.assign(g.qb, g.qb + k);
    const int dim = 1 << k;
    // local position of each operand: qubits[j] <-> matrix index bit k-1-j
    auto local = [&](int q) {
        for (int j = 0; j < k; j++)
            if ((int)qubits[j] == q) return k - 1 - j;
        return -1;
    };
    std::vector<PrimOp> prims;
    lower_gate(g, prims);
    std::vector<cd> M((size_t)dim * dim, 0.0);
    for (int i = 0; i < dim; i++) M[(size_t)i * dim + i] = 1.0;
    bool linear = true;
    for (const PrimOp& p : prims) {
        if (p.kind == PK_C2) {
            // not produced by named gates
            linear = false;
            break;
        }
        cd m[4];
        if (!prim_matrix2(p, m)) {
            linear = false;
            break;
        }
        const int bt = local(p.q0);
        const int bc = p.q1 >= 0 ? local(p.q1) : -1;
        for (int col = 0; col < dim; col++) {
            for (int i0 = 0; i0 < dim; i0++) {
                if (i0 & (1 << bt)) continue;
                if (bc >= 0 && !(i0 & (1 << bc))) continue;
                const int i1 = i0 | (1 << bt);
                const cd a0 = M[(size_t)i0 * dim + col], a1 = M[(size_t)i1 * dim + col];
                M[(size_t)i0 * dim + col] = m[0] * a0 + m[1] * a1;
                M[(size_t)i1 * dim + col] = m[2] * a0 + m[3] * a1;
            }
        }
    }
    if (!linear) {
        k = 0;
        re.clear();
        im.clear();
        return;
    }
    re.resize(M.size());
    im.resize(M.size());
    for (size_t i = 0; i < M.size(); i++) {
        re[i] = M[i].real();
        im[i] = M[i].imag();
    }
}

}  // namespace dmx
