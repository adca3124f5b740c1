// femhip_geom.h -- P1 element geometry shared by the assembly and the
// adjoint's diffusivity gradient.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace femhip {

struct Geo {
  double g[3][3];  // surface gradients of phi_0..phi_2
  double area, jdet, nrm[3];
};

// FEMContext.cpp:73-95: n = normalize(cross(b-a, c-a)), area = |cross|/2,
// J = [b-a, c-a, n], gradients = rows of J^{-1} applied to the reference
// gradients {(-1,-1,0),(1,0,0),(0,1,0)}; J^{-1} by cofactors.
__device__ inline void elem_geom(const float* xyz, const uint32_t* t, Geo& G) {
  double a[3], e1[3], e2[3];
  for (int k = 0; k < 3; k++) {
    a[k] = (double)xyz[3 * (size_t)t[0] + k];
    e1[k] = (double)xyz[3 * (size_t)t[1] + k] - a[k];
    e2[k] = (double)xyz[3 * (size_t)t[2] + k] - a[k];
  }
  double cr0 = e1[1] * e2[2] - e1[2] * e2[1];
  double cr1 = e1[2] * e2[0] - e1[0] * e2[2];
  double cr2 = e1[0] * e2[1] - e1[1] * e2[0];
  double len = sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
  G.nrm[0] = cr0 / len;
  G.nrm[1] = cr1 / len;
  G.nrm[2] = cr2 / len;
  G.area = 0.5 * len;
  G.jdet = len;
  double m00 = e1[0], m01 = e2[0], m02 = G.nrm[0];
  double m10 = e1[1], m11 = e2[1], m12 = G.nrm[1];
  double m20 = e1[2], m21 = e2[2], m22 = G.nrm[2];
  double c00 = m11 * m22 - m12 * m21;
  double c01 = m12 * m20 - m10 * m22;
  double c02 = m10 * m21 - m11 * m20;
  double det = m00 * c00 + m01 * c01 + m02 * c02;
  G.g[1][0] = c00 / det;
  G.g[1][1] = (m02 * m21 - m01 * m22) / det;
  G.g[1][2] = (m01 * m12 - m02 * m11) / det;
  G.g[2][0] = c01 / det;
  G.g[2][1] = (m00 * m22 - m02 * m20) / det;
  G.g[2][2] = (m02 * m10 - m00 * m12) / det;
  for (int k = 0; k < 3; k++) G.g[0][k] = -G.g[1][k] - G.g[2][k];
}

// Unit-diffusivity element stiffness area * grad(phi_a) . grad(phi_b)
// (FEMContext.cpp:102-104).
__device__ inline double elem_kab(const Geo& G, int a, int b) {
  return G.area * (G.g[a][0] * G.g[b][0] + G.g[a][1] * G.g[b][1] + G.g[a][2] * G.g[b][2]);
}

}  // namespace femhip
