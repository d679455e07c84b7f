// Device element math for the 'user potential' models on Q1 quads, Hex8
// bricks and 2-node boundary lines.  An element is processed by a team of
// NPE lanes: lane t first evaluates Gauss point q = t (NQ == NPE for the
// 2-point tensor rules) into LDS, then produces the residual / tangent rows
// of local node a = t.  All geometry and kinematics are FP64; tangent rows
// are accumulated in T (double or float).
//
// Reference: shape functions autopdex/spaces.py:1913-1922 (Q1) and
// :12057-12081 (Hex8); physical gradient dF/dxi . (dX/dxi)^-1
// (spaces.py:12162-12191); potential sum_q w_q det(J_q) psi with the signed
// determinant (models.py:1188-1269); Gauss rule seeder.py:370-385,
// :1046-1062 (xi_0 fastest, weights 1); Neo-Hooke models.py:1122-1146.
#pragma once
#include "plan.hpp"

namespace apx {

// 2-point Gauss-Legendre abscissae, 2*p - 1 as in seeder.gauss_legendre_nd.
constexpr double kGaussA = 2.0 * 0.21132486540518713 - 1.0;
constexpr double kGaussB = 2.0 * 0.7886751345948129 - 1.0;
__device__ __forceinline__ double gauss_x(int i) { return i ? kGaussB : kGaussA; }

// 3-point Gauss-Legendre rule on [-1, 1]: 2 p - 1 and 2 w of the order-4|5 rule
// of seeder.gauss_legendre_1d (seeder.py:387-393), digits as the reference holds them.
constexpr double kGauss3A = 2.0 * 0.1127016653792583 - 1.0;
constexpr double kGauss3C = 2.0 * 0.8872983346207417 - 1.0;
constexpr double kGauss3WE = 2.0 * 0.2777777777777777, kGauss3WM = 2.0 * 0.44444444444444453;
__device__ __forceinline__ double gauss3_x(int i) { return i == 0 ? kGauss3A : (i == 1 ? 0.0 : kGauss3C); }
__device__ __forceinline__ double gauss3_w(int i) { return i == 1 ? kGauss3WM : kGauss3WE; }

// 1D quadratic Lagrange polynomials on the nodes -1, 0, 1 and their derivatives.
__device__ __forceinline__ void lagrange3(double x, double L[3], double dL[3]) {
  L[0] = 0.5 * x * (x - 1.0);
  L[1] = 1.0 - x * x;
  L[2] = 0.5 * x * (x + 1.0);
  dL[0] = x - 0.5;
  dL[1] = -2.0 * x;
  dL[2] = x + 0.5;
}

// ---- shape functions --------------------------------------------------------
template <int NPE, int DIM>
struct Shape;

template <>
struct Shape<4, 2> {  // Q1, nodes (-1,-1), (1,-1), (1,1), (-1,1)
  static constexpr int NQ = 4;
  __device__ static double weight(int) { return 1.0; }  // 2-point rule: weights 2 * 0.5
  __device__ static void eval(int q, double N[4], double dN[4][2]) {
    const double x0 = 0.5 * gauss_x(q & 1), x2 = 0.5 * gauss_x(q >> 1);
    const double x1 = 0.5 - x0, x3 = 0.5 - x2, x4 = x0 + 0.5, x5 = x2 + 0.5;
    N[0] = x1 * x3; N[1] = x3 * x4; N[2] = x4 * x5; N[3] = x1 * x5;
    dN[0][0] = -0.5 * x3; dN[0][1] = -0.5 * x1;
    dN[1][0] = 0.5 * x3;  dN[1][1] = -0.5 * x4;
    dN[2][0] = 0.5 * x5;  dN[2][1] = 0.5 * x4;
    dN[3][0] = -0.5 * x5; dN[3][1] = 0.5 * x1;
  }
};

template <>
struct Shape<8, 3> {  // Hex8, node signs per spaces.py:12057-12081
  static constexpr int NQ = 8;
  __device__ static double weight(int) { return 1.0; }
  __device__ static void eval(int q, double N[8], double dN[8][3]) {
    const double xi[3] = {gauss_x(q & 1), gauss_x((q >> 1) & 1), gauss_x(q >> 2)};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const double s0 = ((a + 1) >> 1) & 1 ? 1.0 : -1.0;  // -,+,+,-,-,+,+,-
      const double s1 = (a >> 1) & 1 ? 1.0 : -1.0;        // -,-,+,+,-,-,+,+
      const double s2 = (a >> 2) & 1 ? 1.0 : -1.0;
      const double f0 = 1.0 + s0 * xi[0], f1 = 1.0 + s1 * xi[1], f2 = 1.0 + s2 * xi[2];
      N[a] = f0 * f1 * f2 * 0.125;
      dN[a][0] = s0 * f1 * f2 * 0.125;
      dN[a][1] = f0 * s1 * f2 * 0.125;
      dN[a][2] = f0 * f1 * s2 * 0.125;
    }
  }
};

// Q2 (9-node) quad of the quadrilateral branch of spaces.fem_iso_line_quad_brick:
// tensor products L_i(xi_0) L_j(xi_1) of the quadratic Lagrange polynomials in the
// node order corners (-1,-1), (1,-1), (1,1), (-1,1), midsides (0,-1), (1,0), (0,1),
// (-1,0), centre (0,0).  3 x 3 Gauss rule, xi_0 fastest (seeder.py:1046-1062).
template <>
struct Shape<9, 2> {
  static constexpr int NQ = 9;
  __device__ static double weight(int q) { return gauss3_w(q % 3) * gauss3_w(q / 3); }
  __device__ static void eval(int q, double N[9], double dN[9][2]) {
    double Lx[3], Ly[3], dLx[3], dLy[3];
    lagrange3(gauss3_x(q % 3), Lx, dLx);
    lagrange3(gauss3_x(q / 3), Ly, dLy);
    constexpr int I[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1};  // xi_0 node index of node a
    constexpr int J[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};  // xi_1 node index of node a
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      N[a] = Lx[I[a]] * Ly[J[a]];
      dN[a][0] = dLx[I[a]] * Ly[J[a]];
      dN[a][1] = Lx[I[a]] * dLy[J[a]];
    }
  }
};

// J_ij = sum_a x_ai dN_aj ; returns signed det J and G = dN J^-1.
template <int NPE, int DIM>
__device__ inline double iso_map(const double (*x)[DIM], const double (*dN)[DIM], double (*G)[DIM]) {
  double J[DIM][DIM];
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < NPE; ++a) s += x[a][i] * dN[a][j];
      J[i][j] = s;
    }
  double det, Ji[DIM][DIM];
  if constexpr (DIM == 2) {
    det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double id = 1.0 / det;
    Ji[0][0] = J[1][1] * id; Ji[0][1] = -J[0][1] * id;
    Ji[1][0] = -J[1][0] * id; Ji[1][1] = J[0][0] * id;
  } else {
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    const double c02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    const double c10 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c11 = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    const double c12 = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    const double c20 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double c21 = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    const double c22 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    det = J[0][0] * c00 + J[0][1] * c10 + J[0][2] * c20;
    const double id = 1.0 / det;
    Ji[0][0] = c00 * id; Ji[0][1] = c01 * id; Ji[0][2] = c02 * id;
    Ji[1][0] = c10 * id; Ji[1][1] = c11 * id; Ji[1][2] = c12 * id;
    Ji[2][0] = c20 * id; Ji[2][1] = c21 * id; Ji[2][2] = c22 * id;
  }
#pragma unroll
  for (int a = 0; a < NPE; ++a)
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < DIM; ++j) s += dN[a][j] * Ji[j][k];
      G[a][k] = s;
    }
  return det;
}

// ---- Poisson with the analytic source of short_example.py:23-33 ----------------
// r_a = sum_q w det (G_a . grad u - b N_a), K_ab = sum_q w det G_a . G_b
template <int NPE_, int DIM_>
struct PoissonSource {
  static constexpr int NPE = NPE_, DIM = DIM_, D = 1, NQ = NPE_, NODAL = 0;
  template <class T>
  struct QP {
    T wdet, G[NPE][DIM], grad[DIM], srcN;  // srcN = b at x_q (residual only)
  };
  template <class T>
  __device__ static void qp(int q, const double (*x)[DIM], const double (*u)[1], const double (*)[2],
                            const ModelParams& p, bool want_res, QP<T>& o) {
    double N[NPE], dN[NPE][DIM], G[NPE][DIM];
    Shape<NPE, DIM>::eval(q, N, dN);
    const double det = iso_map<NPE, DIM>(x, dN, G);
    o.wdet = (T)det;  // Gauss weights are all 1
    double xq[DIM] = {};
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      double g = 0.0;
#pragma unroll
      for (int a = 0; a < NPE; ++a) {
        g += u[a][0] * G[a][k];
        xq[k] += N[a] * x[a][k];
      }
      o.grad[k] = (T)g;
#pragma unroll
      for (int a = 0; a < NPE; ++a) o.G[a][k] = (T)G[a][k];
    }
    o.srcN = 0;
    if (want_res) {
      // b = p0 (sin(p1 x.x) - cos(p1 |x - (p2, p3)|^2))
      double xx = 0.0, yy = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        xx += xq[k] * xq[k];
        const double c = xq[k] - (k == 0 ? p.p[2] : p.p[3]);
        yy += c * c;
      }
      o.srcN = (T)(p.p[0] * (sin(p.p[1] * xx) - cos(p.p[1] * yy)));
    }
  }
  __device__ static double nshape(int q, int a) {
    double N[NPE], dN[NPE][DIM];
    Shape<NPE, DIM>::eval(q, N, dN);
    return N[a];
  }
  __device__ static void res_row(int a, const QP<double>* Q, const ModelParams&, double out[1]) {
    double s = 0.0;
    for (int q = 0; q < NQ; ++q) {
      double gg = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) gg += Q[q].G[a][k] * Q[q].grad[k];
      s += Q[q].wdet * (gg - Q[q].srcN * nshape(q, a));
    }
    out[0] = s;
  }
  template <class T>
  __device__ static void tan_rows(int a, const QP<T>* Q, const ModelParams&, T out[1][NPE]) {
#pragma unroll
    for (int b = 0; b < NPE; ++b) out[0][b] = 0;
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int b = 0; b < NPE; ++b) {
        T gg = 0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) gg += Q[q].G[a][k] * Q[q].G[b][k];
        out[0][b] += Q[q].wdet * gg;
      }
  }
  // Gauss point q's contribution to every node (one thread per element)
  __device__ static void res_gp(int q, const QP<double>& Q, const ModelParams&, double out[NPE][1]) {
    double N[NPE], dN[NPE][DIM];
    Shape<NPE, DIM>::eval(q, N, dN);
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
      double gg = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) gg += Q.G[a][k] * Q.grad[k];
      out[a][0] += Q.wdet * (gg - Q.srcN * N[a]);
    }
  }
  template <class T>
  __device__ static void tan_gp(const QP<T>& Q, const ModelParams&, T K[NPE][NPE]) {
#pragma unroll
    for (int a = 0; a < NPE; ++a)
#pragma unroll
      for (int b = 0; b < NPE; ++b) {
        T gg = 0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) gg += Q.G[a][k] * Q.G[b][k];
        K[a][b] += Q.wdet * gg;
      }
  }
};

// ---- Poisson with nodal coefficient k = exp(N.g) and source f = N.f -------------
template <int NPE_, int DIM_>
struct PoissonVarK {
  static constexpr int NPE = NPE_, DIM = DIM_, D = 1, NQ = NPE_, NODAL = 1;
  template <class T>
  struct QP {
    T wdet, kq, fq, G[NPE][DIM], grad[DIM];
  };
  template <class T>
  __device__ static void qp(int q, const double (*x)[DIM], const double (*u)[1], const double (*nod)[2],
                            const ModelParams&, bool, QP<T>& o) {
    double N[NPE], dN[NPE][DIM], G[NPE][DIM];
    Shape<NPE, DIM>::eval(q, N, dN);
    const double det = iso_map<NPE, DIM>(x, dN, G);
    double g = 0.0, f = 0.0;
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
      g += N[a] * nod[a][0];
      f += N[a] * nod[a][1];
    }
    o.wdet = (T)det;
    o.kq = (T)exp(g);
    o.fq = (T)f;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < NPE; ++a) {
        s += u[a][0] * G[a][k];
        o.G[a][k] = (T)G[a][k];
      }
      o.grad[k] = (T)s;
    }
  }
  __device__ static void res_row(int a, const QP<double>* Q, const ModelParams&, double out[1]) {
    double s = 0.0;
    for (int q = 0; q < NQ; ++q) {
      double N[NPE], dN[NPE][DIM];
      Shape<NPE, DIM>::eval(q, N, dN);
      double gg = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) gg += Q[q].G[a][k] * Q[q].grad[k];
      s += Q[q].wdet * (Q[q].kq * gg - Q[q].fq * N[a]);
    }
    out[0] = s;
  }
  template <class T>
  __device__ static void tan_rows(int a, const QP<T>* Q, const ModelParams&, T out[1][NPE]) {
#pragma unroll
    for (int b = 0; b < NPE; ++b) out[0][b] = 0;
    for (int q = 0; q < NQ; ++q) {
      const T w = Q[q].wdet * Q[q].kq;
#pragma unroll
      for (int b = 0; b < NPE; ++b) {
        T gg = 0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) gg += Q[q].G[a][k] * Q[q].G[b][k];
        out[0][b] += w * gg;
      }
    }
  }
  __device__ static void res_gp(int q, const QP<double>& Q, const ModelParams&, double out[NPE][1]) {
    double N[NPE], dN[NPE][DIM];
    Shape<NPE, DIM>::eval(q, N, dN);
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
      double gg = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) gg += Q.G[a][k] * Q.grad[k];
      out[a][0] += Q.wdet * (Q.kq * gg - Q.fq * N[a]);
    }
  }
  template <class T>
  __device__ static void tan_gp(const QP<T>& Q, const ModelParams&, T K[NPE][NPE]) {
    const T w = Q.wdet * Q.kq;
#pragma unroll
    for (int a = 0; a < NPE; ++a)
#pragma unroll
      for (int b = 0; b < NPE; ++b) {
        T gg = 0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) gg += Q.G[a][k] * Q.G[b][k];
        K[a][b] += w * gg;
      }
  }
};

// ---- Ciarlet Neo-Hooke, psi(F), F = I + grad u (plane strain for DIM = 2) --------
// P = mu F - c1 F^-T, c1 = mu - lam/2 (J^2 - 1);
// K[ai,bk] = sum_q w det (mu d_ik G_a.G_b + c1 H_ak H_bi + c2 H_ai H_bk),
// H_a = F^-T G_a, c2 = lam J^2.
template <int NPE_, int DIM_>
struct NeoHooke {
  static constexpr int NPE = NPE_, DIM = DIM_, D = DIM_, NQ = NPE_, NODAL = 0;
  template <class T>
  struct QP {
    T wdet, c1, c2, G[NPE][DIM], H[NPE][DIM], FG[NPE][DIM];  // FG_a = F G_a (residual)
  };
  template <class T>
  __device__ static void qp(int q, const double (*x)[DIM], const double (*u)[DIM], const double (*)[2],
                            const ModelParams& p, bool want_res, QP<T>& o) {
    const double lam = p.p[0], mu = p.p[1];
    double N[NPE], dN[NPE][DIM], G[NPE][DIM];
    Shape<NPE, DIM>::eval(q, N, dN);
    const double det = iso_map<NPE, DIM>(x, dN, G);
    double F[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double s = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int a = 0; a < NPE; ++a) s += u[a][i] * G[a][j];
        F[i][j] = s;
      }
    double J, Fi[DIM][DIM];
    if constexpr (DIM == 2) {
      J = F[0][0] * F[1][1] - F[0][1] * F[1][0];
      const double id = 1.0 / J;
      Fi[0][0] = F[1][1] * id; Fi[0][1] = -F[0][1] * id;
      Fi[1][0] = -F[1][0] * id; Fi[1][1] = F[0][0] * id;
    } else {
      const double c00 = F[1][1] * F[2][2] - F[1][2] * F[2][1];
      const double c01 = F[0][2] * F[2][1] - F[0][1] * F[2][2];
      const double c02 = F[0][1] * F[1][2] - F[0][2] * F[1][1];
      const double c10 = F[1][2] * F[2][0] - F[1][0] * F[2][2];
      const double c11 = F[0][0] * F[2][2] - F[0][2] * F[2][0];
      const double c12 = F[0][2] * F[1][0] - F[0][0] * F[1][2];
      const double c20 = F[1][0] * F[2][1] - F[1][1] * F[2][0];
      const double c21 = F[0][1] * F[2][0] - F[0][0] * F[2][1];
      const double c22 = F[0][0] * F[1][1] - F[0][1] * F[1][0];
      J = F[0][0] * c00 + F[0][1] * c10 + F[0][2] * c20;
      const double id = 1.0 / J;
      Fi[0][0] = c00 * id; Fi[0][1] = c01 * id; Fi[0][2] = c02 * id;
      Fi[1][0] = c10 * id; Fi[1][1] = c11 * id; Fi[1][2] = c12 * id;
      Fi[2][0] = c20 * id; Fi[2][1] = c21 * id; Fi[2][2] = c22 * id;
    }
    // The reference differentiates psi with jacrev, whose derivative of log J
    // is 1/J for any J != 0; the closed forms below are that same derivative.
    const double c1 = mu - 0.5 * lam * (J * J - 1.0);
    o.wdet = (T)(Shape<NPE, DIM>::weight(q) * det);
    o.c1 = (T)c1;
    o.c2 = (T)(lam * J * J);
#pragma unroll
    for (int a = 0; a < NPE; ++a)
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        double h = 0.0, fg = 0.0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
          h += Fi[j][k] * G[a][j];
          fg += F[k][j] * G[a][j];
        }
        o.G[a][k] = (T)G[a][k];
        o.H[a][k] = (T)h;
        o.FG[a][k] = (T)(want_res ? mu * fg : 0.0);
      }
  }
  __device__ static void res_gp(int, const QP<double>& Q, const ModelParams&, double out[NPE][DIM]) {
#pragma unroll
    for (int a = 0; a < NPE; ++a)
#pragma unroll
      for (int i = 0; i < DIM; ++i) out[a][i] += Q.wdet * (Q.FG[a][i] - Q.c1 * Q.H[a][i]);
  }
  // r_ai = sum_q w det (mu (F G_a)_i - c1 H_ai)
  __device__ static void res_row(int a, const QP<double>* Q, const ModelParams&, double out[DIM]) {
#pragma unroll
    for (int i = 0; i < DIM; ++i) out[i] = 0.0;
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int i = 0; i < DIM; ++i) out[i] += Q[q].wdet * (Q[q].FG[a][i] - Q[q].c1 * Q[q].H[a][i]);
  }
  template <class T>
  __device__ static void tan_rows(int a, const QP<T>* Q, const ModelParams& p, T out[DIM][NPE * DIM]) {
    const T mu = (T)p.p[1];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int B = 0; B < NPE * DIM; ++B) out[i][B] = 0;
    for (int q = 0; q < NQ; ++q) {
      const T w = Q[q].wdet, c1 = Q[q].c1 * w, c2 = Q[q].c2 * w;
      T Ha[DIM], Ga[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        Ha[k] = Q[q].H[a][k];
        Ga[k] = Q[q].G[a][k];
      }
#pragma unroll
      for (int b = 0; b < NPE; ++b) {
        T gg = 0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) gg += Ga[k] * Q[q].G[b][k];
        gg *= mu * w;
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
          for (int k = 0; k < DIM; ++k)
            out[i][b * DIM + k] += (i == k ? gg : T(0)) + c1 * Ha[k] * Q[q].H[b][i] + c2 * Ha[i] * Q[q].H[b][k];
      }
    }
  }
};

// ---- dead-load traction on 2-node lines in 2D ------------------------------------
// r_ai = - sum_q w_q |dX/dxi| N_a(xi_q) t_i, zero tangent.
struct LineTraction {
  static constexpr int NPE = 2, DIM = 2, D = 2, NQ = 2, NODAL = 0;
  template <class T>
  struct QP {
    T wlen;
  };
  template <class T>
  __device__ static void qp(int q, const double (*x)[2], const double (*)[2], const double (*)[2],
                            const ModelParams&, bool, QP<T>& o) {
    const double tx = 0.5 * (x[1][0] - x[0][0]), ty = 0.5 * (x[1][1] - x[0][1]);
    o.wlen = (T)sqrt(tx * tx + ty * ty);  // weight 1
  }
  __device__ static void res_row(int a, const QP<double>* Q, const ModelParams& p, double out[2]) {
    out[0] = out[1] = 0.0;
    for (int q = 0; q < 2; ++q) {
      const double xi = gauss_x(q);
      const double Na = a == 0 ? 0.5 * (1.0 - xi) : 0.5 * (1.0 + xi);
      out[0] -= Q[q].wlen * Na * p.p[0];
      out[1] -= Q[q].wlen * Na * p.p[1];
    }
  }
  template <class T>
  __device__ static void tan_rows(int, const QP<T>*, const ModelParams&, T out[2][4]) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int B = 0; B < 4; ++B) out[i][B] = 0;
  }
};

// ---- dead-load traction on 3-node (quadratic) lines in 2D ----------------------------
// Nodes at xi = -1, 1, 0 (line branch of spaces.fem_iso_line_quad_brick), 3-point rule:
// r_ai = - sum_q w_q |dX/dxi(xi_q)| N_a(xi_q) t_i, zero tangent -- the residual of
// isoparametric_surface_element_galerkin (models.py:1723-1850) with neumann_weak (:744-780).
struct LineTraction3 {
  static constexpr int NPE = 3, DIM = 2, D = 2, NQ = 3, NODAL = 0;
  template <class T>
  struct QP {
    T wlen;
  };
  __device__ static void shape(double xi, double N[3], double dN[3]) {
    N[0] = 0.5 * xi * (xi - 1.0);
    N[1] = 0.5 * xi * (xi + 1.0);
    N[2] = 1.0 - xi * xi;
    dN[0] = xi - 0.5;
    dN[1] = xi + 0.5;
    dN[2] = -2.0 * xi;
  }
  template <class T>
  __device__ static void qp(int q, const double (*x)[2], const double (*)[2], const double (*)[2],
                            const ModelParams&, bool, QP<T>& o) {
    double N[3], dN[3];
    shape(gauss3_x(q), N, dN);
    double tx = 0.0, ty = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      tx += x[a][0] * dN[a];
      ty += x[a][1] * dN[a];
    }
    o.wlen = (T)(gauss3_w(q) * sqrt(tx * tx + ty * ty));
  }
  __device__ static void res_row(int a, const QP<double>* Q, const ModelParams& p, double out[2]) {
    out[0] = out[1] = 0.0;
    for (int q = 0; q < 3; ++q) {
      double N[3], dN[3];
      shape(gauss3_x(q), N, dN);
      out[0] -= Q[q].wlen * N[a] * p.p[0];
      out[1] -= Q[q].wlen * N[a] * p.p[1];
    }
  }
  template <class T>
  __device__ static void tan_rows(int, const QP<T>*, const ModelParams&, T out[2][6]) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int B = 0; B < 6; ++B) out[i][B] = 0;
  }
};

}  // namespace apx
