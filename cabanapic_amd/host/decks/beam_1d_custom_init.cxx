// Deck-override smoke test (the role of the reference's decks/custom_init.cxx
// in tests/decks): a deck that replaces the particle initializer.  Two cold
// counter-streaming beams along x in 32 x 1 x 1 cells, 100 particles per
// cell, offsets spread along x, 30 steps.  Passes when the run exits 0.
#include <cmath>

#include "../deck.h"

class Beams_Along_X : public Particle_Initializer {
 public:
  void init(particle_list_t& particles, size_t nx, size_t ny, size_t nz, size_t ng, real_ dxp,
            size_t nppc, real_ w, real_ v0, real_, real_, real_) override {
    std::cout << "Using a custom particle initializer" << std::endl;
    auto px = slice<PositionX>(particles);
    auto py = slice<PositionY>(particles);
    auto pz = slice<PositionZ>(particles);
    auto vx = slice<VelocityX>(particles);
    auto vy = slice<VelocityY>(particles);
    auto vz = slice<VelocityZ>(particles);
    auto wt = slice<Weight>(particles);
    auto cell = cell_slice(particles);
    const real_ gam = 1.0 / std::sqrt(1.0 - v0 * v0);
    simd_for_each(particles, [&](size_t s, size_t i) {
      const size_t idx = s * particle_list_t::vector_length + i;
      const size_t pair = idx / 2;
      const double sign = (idx % 2 == 0) ? 1.0 : -1.0;
      const size_t k = (2 * pair) % nppc;
      const size_t ix = 2 * pair / nppc + 1;
      const real_ x = k * dxp + 0.5 * dxp - 1.0;
      px.access(s, i) = x;
      py.access(s, i) = 0.0;
      pz.access(s, i) = 0.0;
      wt.access(s, i) = w;
      cell.access(s, i) = (int32_t)VOXEL(ix, 1, 1, nx, ny, nz, ng);
      const double pert = 0.0001 * std::sin(2.0 * M_PI * ((x + 1.0 + ix * 2.0) / (2.0 * nx)));
      vx.access(s, i) = sign * v0 * gam * (1.0 + pert);
      vy.access(s, i) = 0;
      vz.access(s, i) = 0;
    });
  }
};

Input_Deck::Input_Deck() {
  particle_initer = new Beams_Along_X();
  nx = 32;
  ny = 1;
  nz = 1;
  num_steps = 30;
  nppc = 100;
  v0 = 0.0866025403784439;
  const real_ gam = 1.0 / std::sqrt(1.0 - v0 * v0);
  len_x_global = 6.28318530717959 * (gam * std::sqrt((double)gam));
  Npe = n0 * len_x_global * len_y_global * len_z_global;
  dt = 0.99 * courant_length(len_x_global, len_y_global, len_z_global, nx, ny, nz) / c;
  n0 = 2.0;
}
