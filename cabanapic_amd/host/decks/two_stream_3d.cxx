// 3D two-stream: 64^3 periodic cells, 64 particles per cell in two
// counter-streaming electron beams at +-0.2 c along x (drift momentum
// 0.2/sqrt(1-0.04)), thermal spread u_th = 0.01 c per component, seed 7,
// drift modulated by 1 + 1e-3 sin(2 pi x / Lx), dt = 0.99 Courant, 200 steps.
#include <cmath>

#include "../plasma_init.h"

Input_Deck::Input_Deck() {
  auto* p = new Plasma_Initializer();
  p->seed = 7;
  p->uth[0] = p->uth[1] = p->uth[2] = 0.01;
  p->drift[0] = 0.2 / std::sqrt(1.0 - 0.04);
  p->pert_amp = 1e-3;
  p->pert_mode = 1;
  p->counter_streaming = true;
  particle_initer = p;
  nx = ny = nz = 64;
  nppc = 64;
  num_steps = 200;
  v0 = 0.2;
  len_x_global = len_y_global = len_z_global = (real_)(2.0 * M_PI * 0.2 / 0.5);
  dt = 0.99 * courant_length(len_x_global, len_y_global, len_z_global, nx, ny, nz) / c;
  n0 = 1.0;
}
