// 1D-periodic two-stream in x: 32 x 1 x 1 cells, Lx = 2 pi 0.2 / 0.5, two
// cold counter-streaming electron beams at +-0.2 c along x (drift momentum
// 0.2/sqrt(1-0.04)), 32 particles per cell per beam (alternating), offsets
// uniform-random from seed 1234, drift modulated by 1 + 1e-3 sin(2 pi x / Lx),
// immobile ions, EM solver, dt = 0.99 Courant, 1000 steps.  This is the
// reference's decks/2stream-short.cxx case with the beams laid out along x.
#include <cmath>

#include "../plasma_init.h"

Input_Deck::Input_Deck() {
  auto* p = new Plasma_Initializer();
  p->seed = 1234;
  p->drift[0] = 0.2 / std::sqrt(1.0 - 0.04);
  p->pert_amp = 1e-3;
  p->pert_mode = 1;
  p->counter_streaming = true;
  particle_initer = p;
  nx = 32;
  ny = nz = 1;
  nppc = 64;
  num_steps = 1000;
  v0 = 0.2;
  len_x_global = (real_)(2.0 * M_PI * 0.2 / 0.5);
  dt = 0.99 * courant_length(len_x_global, len_y_global, len_z_global, nx, ny, nz) / c;
  n0 = 1.0;
}
