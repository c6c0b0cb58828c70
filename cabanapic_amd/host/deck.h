// deck.h -- the input-deck interface of the reference (src/input/deck.h),
// kept at the source level so that a deck written against it compiles here:
// a deck translation unit defines Input_Deck::Input_Deck(), may install its
// own Particle_Initializer / Field_Initializer / Run_Finalizer, and sets the
// public parameters below.  derive_params, courant_length and the default
// initializers follow the reference's definitions (with its literal types,
// so the initial state is bit-identical to the reference's).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <iostream>

#include "pic_types.h"

// src/input/deck.h:416-417 (decks stringify build-time paths with them)
#ifndef STRINGIFY
#define STRINGIFY(s) #s
#endif
#ifndef EXPAND
#define EXPAND(s) STRINGIFY(s)
#endif

// Hook run after the simulation loop (src/input/deck.h:15-21).
class Run_Finalizer {
 public:
  virtual ~Run_Finalizer() = default;
  virtual void finalize() {}
};

// Initial fields; the default zeroes E and B (src/input/deck.h:23-66).
class Field_Initializer {
 public:
  using real_ = real_t;
  virtual ~Field_Initializer() = default;
  virtual void init(field_array_t& fields, size_t, size_t, size_t, size_t, real_, real_, real_,
                    real_, real_, real_) {
    std::cout << "Default field init" << std::endl;
    for (int m = FIELD_EX; m <= FIELD_CBZ; ++m) {
      real_t* a = fields.member(m);
      for (size_t i = 0; i < fields.size(); ++i) a[i] = 0.0;
    }
  }
};

// Initial particles; the default is the reference's two counter-streaming
// beams along x with the offset spread along y (src/input/deck.h:79-158).
class Particle_Initializer {
 public:
  using real_ = real_t;
  virtual ~Particle_Initializer() = default;
  virtual void init(particle_list_t& particles, size_t nx, size_t ny, size_t, size_t, real_ dxp,
                    size_t nppc, real_ w, real_ v0, real_, real_, real_) {
    std::cout << "Default particle init" << std::endl;
    auto px = slice<PositionX>(particles);
    auto py = slice<PositionY>(particles);
    auto pz = slice<PositionZ>(particles);
    auto vx = slice<VelocityX>(particles);
    auto vy = slice<VelocityY>(particles);
    auto vz = slice<VelocityZ>(particles);
    auto wt = slice<Weight>(particles);
    auto cell = cell_slice(particles);
    simd_for_each(particles, [&](size_t s, size_t i) {
      const size_t idx = s * particle_list_t::vector_length + i;
      const size_t pair = idx / 2;
      const int sign = (idx % 2 == 0) ? 1 : -1;
      const int pic = (int)((2 * pair) % nppc);  // both members of a pair share it
      const real_ y = pic * dxp + 0.5 * dxp - 1.0;
      const int pre_ghost = (int)(2 * pair / nppc);
      px.access(s, i) = 0.0;
      py.access(s, i) = y;
      pz.access(s, i) = 0.0;
      wt.access(s, i) = w;
      cell.access(s, i) = (int32_t)(pre_ghost * (nx + 2) + (nx + 2) * (ny + 2) + (nx + 2) + 1);
      const real_ gam = 1.0 / std::sqrt(1.0 - v0 * v0);
      const real_t na = 0.0001 * std::sin(2.0 * 3.1415926 * ((y + 1.0 + pre_ghost * 2) / (double)(2 * ny)));
      vx.access(s, i) = sign * v0 * gam * (1.0 + na * sign);
      vy.access(s, i) = 0;
      vz.access(s, i) = 0;
    });
  }
};

class _Input_Deck {
 public:
  using real_ = real_t;

  // user hooks (owned by the deck for its whole life, as in the reference)
  Particle_Initializer* particle_initer;
  Field_Initializer* field_initer;
  Run_Finalizer* run_finalizer;

  _Input_Deck()
      : particle_initer(new Particle_Initializer),
        field_initer(new Field_Initializer),
        run_finalizer(new Run_Finalizer) {}

  // sqrt(1 / sum over non-degenerate axes of (n/L)^2)  (src/input/deck.h:191-198)
  static real_ courant_length(real_ lx, real_ ly, real_ lz, size_t nx, size_t ny, size_t nz) {
    real_ w0, w1 = 0;
    if (nx > 1) w0 = nx / lx, w1 += w0 * w0;
    if (ny > 1) w0 = ny / ly, w1 += w0 * w0;
    if (nz > 1) w0 = nz / lz, w1 += w0 * w0;
    return std::sqrt((double)(1 / w1));
  }

  void finalize() { run_finalizer->finalize(); }

  void initialize_particles(particle_list_t& particles, size_t nx, size_t ny, size_t nz, size_t ng,
                            real_ dxp, size_t nppc, real_ w, real_ v0) {
    particle_initer->init(particles, nx, ny, nz, ng, dxp, nppc, w, v0, len_x_global, len_y_global,
                          len_z_global);
  }

  void initialize_fields(field_array_t& fields, size_t nx, size_t ny, size_t nz, size_t ng,
                         real_ Lx, real_ Ly, real_ Lz, real_ dx, real_ dy, real_ dz) {
    field_initer->init(fields, nx, ny, nz, ng, Lx, Ly, Lz, dx, dy, dz);
  }

  // normalised units (src/input/deck.h:254-266)
  real_ de = 1.0;
  real_ ec = 1.0;
  real_ me = 1.0;
  real_ mu = 1.0;
  real_ c = 1.0;
  real_ eps = 1.0;
  real_ qsp = -ec;

  // run parameters (src/input/deck.h:268-295)
  real_ n0 = 1.0;
  size_t num_species = 1;
  size_t nx = 16;
  size_t ny = 1;
  size_t nz = 1;
  size_t num_ghosts = 1;
  size_t nppc = 1;
  real_ dt = 1.0;
  int num_steps = 2;
  real_ len_x_global = 1.0;
  real_ len_y_global = 1.0;
  real_ len_z_global = 1.0;
  real_t Npe = -1;
  real_t Ne = -1;
  real_ v0 = 1.0;
  Boundary BOUNDARY_TYPE = Boundary::Periodic;

  // derived (src/input/deck.h:297-309)
  real_ dx, dy, dz;
  real_ len_x, len_y, len_z;
  size_t num_cells;  // including ghosts
  long num_particles = -1;
  bool perform_uncenter = false;

  void print_run_details() {
    std::cout << "#~~~ Run Specifications ~~~ " << std::endl;
    std::cout << "#Nx: " << nx << " Ny: " << ny << " Nz: " << nz
              << " Num Ghosts: " << num_ghosts << ". Cells Total: " << num_cells << std::endl;
    std::cout << "#Len X: " << len_x << " Len Y: " << len_y << " Len Z: " << len_z << std::endl;
    std::cout << "#Approx Particle Count: " << num_particles << " (nppc: " << nppc << ")"
              << std::endl;
    std::cout << "#~~~~~~~~~~~~~~~~~~~~~~~~~~ " << std::endl << std::endl;
  }

  // src/input/deck.h:323-348
  void derive_params() {
    len_x = len_x_global;
    len_y = len_y_global;
    len_z = len_z_global;
    dx = len_x / nx;
    dy = len_y / ny;
    dz = len_z / nz;
    num_cells = (nx + 2 * num_ghosts) * (ny + 2 * num_ghosts) * (nz + 2 * num_ghosts);
    if (num_particles < 0) {
      num_particles = (long)(nx * ny * nz * nppc);
      if (Ne < 0) Ne = num_particles;
    }
    if (Npe < 0) Npe = n0 * len_x_global * len_y_global * len_z_global;
  }
};

// A deck translation unit defines the constructor (src/input/deck.h:415-429);
// without one, the default deck is the 1D-in-y electromagnetic two-stream
// setup the reference ships as its default.
class Input_Deck : public _Input_Deck {
 public:
  Input_Deck();
};

extern Input_Deck deck;
