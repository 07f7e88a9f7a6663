// Surface constraints as inline device functions, FP64, shared by the
// single-system integrators and the replica ensembles.
//
//   sphere           src/constraints/constraint_sphere.cpp:38-81
//   plane            src/constraints/constraint_plane.cpp:36-160
//   none             src/constraints/constraint_none.hpp:58-80
//   cylinder         src/constraints/constraint_cylinder.cpp:30-66, constraint_cylinder.hpp:55-67
//   torus            src/constraints/constraint_torus.cpp:26-52 (generic SHAKE projection)
//   ellipsoid        src/constraints/constraint_ellipsoid.cpp:26-48 (generic SHAKE projection)
//   SHAKE projection src/constraints/constraint.cpp:40-87
//   rotate/project   src/constraints/constraint.cpp:96-173
//   periodic wrap    src/system/system.cpp:1119-1155
//
// Each function takes a template bitmask SET of the constraint kinds it is
// compiled for (bit k = kind k), so kernels that only ever see some kinds
// (the planar fused step, the replica ensembles) carry no code for the others.
#pragma once
#include "samos_device.hpp"

namespace samos {

// One particle's kinematic state in registers; dix/diy/diz collect image-flag
// changes from periodic wraps.
struct PS {
  double x, y, z, vx, vy, vz, nx, ny, nz;
  int dix, diy, diz;
};


// System::enforce_periodic (src/system/system.cpp:1119-1155)
__device__ __forceinline__ void enforce_periodic(const BoxD& b, PS& p) {
  if (!b.periodic) return;
  if (p.x <= b.xlo) { p.x += b.Lx; p.dix--; } else if (p.x > b.xhi) { p.x -= b.Lx; p.dix++; }
  if (p.y <= b.ylo) { p.y += b.Ly; p.diy--; } else if (p.y > b.yhi) { p.y -= b.Ly; p.diy++; }
  if (p.z <= b.zlo) { p.z += b.Lz; p.diz--; } else if (p.z > b.zhi) { p.z -= b.Lz; p.diz++; }
}

constexpr int kSurfBasic = 0x07;   // none, plane, sphere
constexpr int kSurfAll = 0x3f;     // and cylinder, torus, ellipsoid
constexpr int kSurfPlane = 0x02;

// Implicit function g(x) = 0 of the torus (kind 4) or the ellipsoid (kind 5)
// and its gradient, in the reference's operation order.
__device__ __forceinline__ double surf_value(const ConstraintD& c, double x, double y, double z) {
  if (c.kind == 4) {
    const double ft = c.q0 - sqrt(x * x + y * y);
    return ft * ft + z * z - c.q1 * c.q1;
  }
  return x * x / c.q0 / c.q0 + y * y / c.q1 / c.q1 + z * z / c.q2 / c.q2 - 1.0;
}

__device__ __forceinline__ void surf_gradient(const ConstraintD& c, double x, double y, double z, double& gx,
                                              double& gy, double& gz) {
  if (c.kind == 4) {
    const double sxy = sqrt(x * x + y * y);
    const double fact = -2.0 * (c.q0 - sxy) / sxy;
    gx = x * fact; gy = y * fact; gz = 2.0 * z;
  } else {
    gx = 2.0 * x / c.q0 / c.q0; gy = 2.0 * y / c.q1 / c.q1; gz = 2.0 * z / c.q2 / c.q2;
  }
}

// Unit normal of kinds 2-5 at the particle (compute_normal of each surface).
__device__ __forceinline__ void surf_normal(const ConstraintD& c, const PS& p, double& Nx, double& Ny, double& Nz) {
  if (c.kind == 2) {
    Nx = p.x / c.r; Ny = p.y / c.r; Nz = p.z / c.r;
  } else if (c.kind == 3) {
    Nx = p.x / c.r; Ny = p.y / c.r; Nz = 0.0;
  } else {
    surf_gradient(c, p.x, p.y, p.z, Nx, Ny, Nz);
    const double len = sqrt(Nx * Nx + Ny * Ny + Nz * Nz);
    Nx /= len; Ny /= len; Nz /= len;
  }
}

// Constraint::enforce of the one active constraint; returns true when the
// plane constraint zeroed the particle's f_z.  plane_dirs = false leaves out
// the plane's director normalisation, for a caller that has already
// normalised the same director in this step (on a plane the normal is fixed,
// so a second normalisation of a unit director returns it to within an ulp).
template <int SET = kSurfAll>
__device__ __forceinline__ bool enforce(const BoxD& b, const ConstraintD& c, PS& p, bool plane_dirs = true) {
  if ((SET & 0x01) && c.kind == 0) {
    enforce_periodic(b, p);
  } else if ((SET & 0x02) && c.kind == 1) {
    const double xlo = -0.5 * b.Lx, xhi = 0.5 * b.Lx, ylo = -0.5 * b.Ly, yhi = 0.5 * b.Ly;
    p.z = c.zpos; p.vz = 0.0;
    if (b.periodic) enforce_periodic(b, p);
    else if (!c.unlimited) {
      if (p.x < xlo) { p.x = xlo; p.vx = -p.vx; } else if (p.x > xhi) { p.x = xhi; p.vx = -p.vx; }
      if (p.y < ylo) { p.y = ylo; p.vy = -p.vy; } else if (p.y > yhi) { p.y = yhi; p.vy = -p.vy; }
    }
    p.nz = 0.0;
    if (plane_dirs) {
      const double len_n = sqrt(p.nx * p.nx + p.ny * p.ny);
      double inv_len = 1.0;
      if (len_n != 0.0) inv_len = 1.0 / sqrt(p.nx * p.nx + p.ny * p.ny);
      p.nx *= inv_len; p.ny *= inv_len;
    }
    return true;
  } else if ((SET & 0x04) && c.kind == 2) {
    const double R0 = c.r;
    const double R = sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    const double sc = R0 / R;
    p.x *= sc; p.y *= sc; p.z *= sc;
    const double Nx = p.x / R0, Ny = p.y / R0, Nz = p.z / R0;
    const double vN = p.vx * Nx + p.vy * Ny + p.vz * Nz;
    const double nN = p.nx * Nx + p.ny * Ny + p.nz * Nz;
    p.vx -= vN * Nx; p.vy -= vN * Ny; p.vz -= vN * Nz;
    p.nx -= nN * Nx; p.ny -= nN * Ny; p.nz -= nN * Nz;
    const double inv_len = 1.0 / sqrt(p.nx * p.nx + p.ny * p.ny + p.nz * p.nz);
    p.nx *= inv_len; p.ny *= inv_len; p.nz *= inv_len;
  } else if ((SET & 0x08) && c.kind == 3) {
    // radial scaling in the xy plane; z is wrapped without an image-flag change
    const double R = sqrt(p.x * p.x + p.y * p.y);
    const double sc = c.r / R;
    p.x *= sc; p.y *= sc;
    const double Nx = p.x / c.r, Ny = p.y / c.r;
    const double vN = p.vx * Nx + p.vy * Ny;
    const double nN = p.nx * Nx + p.ny * Ny;
    p.vx -= vN * Nx; p.vy -= vN * Ny;
    p.nx -= nN * Nx; p.ny -= nN * Ny;
    const double inv_len = 1.0 / sqrt(p.nx * p.nx + p.ny * p.ny + p.nz * p.nz);
    p.nx *= inv_len; p.ny *= inv_len; p.nz *= inv_len;
    if (b.periodic) {
      if (p.z > b.zhi) p.z -= b.Lz;
      else if (p.z < b.zlo) p.z += b.Lz;
    }
  } else if ((SET & 0x30) && (c.kind == 4 || c.kind == 5)) {
    // SHAKE: step along the gradient taken at the entry point until |g| <= tol
    double rgx, rgy, rgz;
    surf_gradient(c, p.x, p.y, p.z, rgx, rgy, rgz);
    for (int iter = 0; iter < c.max_iter; ++iter) {
      double gx, gy, gz;
      surf_gradient(c, p.x, p.y, p.z, gx, gy, gz);
      const double sdot = gx * rgx + gy * rgy + gz * rgz;
      const double g = surf_value(c, p.x, p.y, p.z);
      if (!(fabs(g) > c.tol)) break;
      const double lambda = g / sdot;
      p.x -= lambda * rgx; p.y -= lambda * rgy; p.z -= lambda * rgz;
    }
    double Nx, Ny, Nz;
    surf_normal(c, p, Nx, Ny, Nz);
    const double vN = p.vx * Nx + p.vy * Ny + p.vz * Nz;
    const double nN = p.nx * Nx + p.ny * Ny + p.nz * Nz;
    p.vx -= vN * Nx; p.vy -= vN * Ny; p.vz -= vN * Nz;
    p.nx -= nN * Nx; p.ny -= nN * Ny; p.nz -= nN * Nz;
    const double inv_len = 1.0 / sqrt(p.nx * p.nx + p.ny * p.ny + p.nz * p.nz);
    p.nx *= inv_len; p.ny *= inv_len; p.nz *= inv_len;
    enforce_periodic(b, p);
  }
  return false;
}

template <int SET = kSurfAll>
__device__ __forceinline__ double project_torque(const ConstraintD& c, const PS& p, double tx, double ty, double tz) {
  if ((SET & 0x02) && c.kind == 1) return tz;
  if ((SET & 0x3c) && c.kind >= 2) {
    double Nx, Ny, Nz;
    surf_normal(c, p, Nx, Ny, Nz);
    return tx * Nx + ty * Ny + tz * Nz;
  }
  return 0.0;
}

template <int SET = kSurfAll>
__device__ __forceinline__ void rotate_director(const ConstraintD& c, PS& p, double phi) {
  if ((SET & 0x02) && c.kind == 1) {
    double sn, cs;
    sincos(phi, &sn, &cs);
    const double nx = cs * p.nx + (-sn) * p.ny;
    const double ny = sn * p.nx + cs * p.ny;
    const double len = sqrt(nx * nx + ny * ny);
    p.nx = nx / len; p.ny = ny / len;
  } else if ((SET & 0x3c) && c.kind >= 2) {
    double U, V, W;
    surf_normal(c, p, U, V, W);
    double sn, cs;
    sincos(phi, &sn, &cs);
    const double un = U * p.nx + V * p.ny + W * p.nz;
    const double nx = U * un * (1.0 - cs) + p.nx * cs + (-W * p.ny + V * p.nz) * sn;
    const double ny = V * un * (1.0 - cs) + p.ny * cs + (W * p.nx - U * p.nz) * sn;
    const double nz = W * un * (1.0 - cs) + p.nz * cs + (-V * p.nx + U * p.ny) * sn;
    const double len = sqrt(nx * nx + ny * ny + nz * nz);
    p.nx = nx / len; p.ny = ny / len; p.nz = nz / len;
  }
}

template <int SET = kSurfAll>
__device__ __forceinline__ void rotate_velocity(const ConstraintD& c, PS& p, double phi) {
  if ((SET & 0x02) && c.kind == 1) {
    double sn, cs;
    sincos(phi, &sn, &cs);
    const double vx = cs * p.vx + (-sn) * p.vy;
    const double vy = sn * p.vx + cs * p.vy;
    p.vx = vx; p.vy = vy;
  } else if ((SET & 0x3c) && c.kind >= 2) {
    double U, V, W;
    surf_normal(c, p, U, V, W);
    double sn, cs;
    sincos(phi, &sn, &cs);
    const double uv = U * p.vx + V * p.vy + W * p.vz;
    const double vx = U * uv * (1.0 - cs) + p.vx * cs + (-W * p.vy + V * p.vz) * sn;
    const double vy = V * uv * (1.0 - cs) + p.vy * cs + (W * p.vx - U * p.vz) * sn;
    const double vz = W * uv * (1.0 - cs) + p.vz * cs + (-V * p.vx + U * p.vy) * sn;
    p.vx = vx; p.vy = vy; p.vz = vz;
  }
}

}  // namespace samos
