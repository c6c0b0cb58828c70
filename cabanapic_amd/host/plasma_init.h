// plasma_init.h -- a Particle_Initializer for uniform thermal plasmas,
// beams and electron columns, backed by the library's counter-based loader (cpic_init_plasma), so
// a deck and the Python decks of the benchmark produce identical particles.
#pragma once

#include <cstdio>
#include <cstdlib>

#include "../../include/cabanapic_amd.h"
#include "deck.h"

class Plasma_Initializer : public Particle_Initializer {
 public:
  uint64_t seed = 1;
  double uth[3] = {0, 0, 0};
  double drift[3] = {0, 0, 0};
  double pert_amp = 0;
  int pert_mode = 0;
  bool counter_streaming = false;
  // electron column (profile 1 of cpic_plasma_spec): ring radius and width,
  // peak charge density and guide field, in the deck's units
  bool column = false;
  double ring_radius = 0, ring_width = 0, column_charge = 0, guide_bz = 0;

  void init(particle_list_t& particles, size_t nx, size_t ny, size_t nz, size_t, real_, size_t nppc,
            real_ w, real_, real_ Lx, real_ Ly, real_) override {
    cpic_plasma_spec s{};
    s.seed = seed;
    s.nx = (int64_t)nx;
    s.ny = (int64_t)ny;
    s.nz = (int64_t)nz;
    s.nppc = (int64_t)nppc;
    s.weight = w;
    for (int k = 0; k < 3; ++k) {
      s.uth[k] = uth[k];
      s.drift[k] = drift[k];
    }
    s.pert_amp = pert_amp;
    s.len_x = Lx;
    s.pert_mode = pert_mode;
    s.counter_streaming = counter_streaming ? 1 : 0;
    fill_column(s, Lx, Ly);
    const int rc = cpic_init_plasma(&s, (int32_t)sizeof(real_t), (int64_t)particles.size(),
                                    particles.data(PositionX), particles.data(PositionY),
                                    particles.data(PositionZ), particles.data(VelocityX),
                                    particles.data(VelocityY), particles.data(VelocityZ),
                                    particles.data(Weight), particles.cells());
    if (rc != CPIC_OK) {
      std::fprintf(stderr, "cpic_init_plasma failed (%d)\n", rc);
      std::exit(1);
    }
  }

  void fill_column(cpic_plasma_spec& s, double Lx, double Ly) const {
    if (!column) return;
    s.profile = 1;
    s.len_x = Lx;
    s.len_y = Ly;
    s.ring_radius = ring_radius;
    s.ring_width = ring_width;
    s.column_charge = column_charge;
    s.guide_bz = guide_bz;
    s.pert_amp = pert_amp;
    s.pert_mode = pert_mode;
  }
};

// The initial fields of an electron column: its radial E on the Yee edges of
// every voxel (ghosts included) and the guide field in cbz.
class Column_Field_Initializer : public Field_Initializer {
 public:
  const Plasma_Initializer* plasma = nullptr;

  void init(field_array_t& fields, size_t nx, size_t ny, size_t nz, size_t, real_t Lx, real_t Ly,
            real_t, real_t, real_t, real_t) override {
    cpic_plasma_spec s{};
    s.nx = (int64_t)nx;
    s.ny = (int64_t)ny;
    s.nz = (int64_t)nz;
    plasma->fill_column(s, Lx, Ly);
    const int rc = cpic_init_column_fields(&s, (int32_t)sizeof(real_t), fields.data());
    if (rc != CPIC_OK) {
      std::fprintf(stderr, "cpic_init_column_fields failed (%d)\n", rc);
      std::exit(1);
    }
  }
};
