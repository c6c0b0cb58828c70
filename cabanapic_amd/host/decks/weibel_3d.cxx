// 3D Weibel-type anisotropic plasma (the precision / throughput
// configuration): 128^3 periodic cells, 32 particles per cell (67,108,864
// electrons), bi-Maxwellian momenta u_th = (0.3, 0.05, 0.05) c, uniform
// positions from seed 11, immobile neutralising ions, dx = 0.1 c/w_pe,
// dt = 0.99 Courant, 500 steps.
#include "../plasma_init.h"

Input_Deck::Input_Deck() {
  auto* p = new Plasma_Initializer();
  p->seed = 11;
  p->uth[0] = 0.3;
  p->uth[1] = 0.05;
  p->uth[2] = 0.05;
  particle_initer = p;
  nx = ny = nz = 128;
  nppc = 32;
  num_steps = 500;
  len_x_global = len_y_global = len_z_global = (real_)(128 * 0.1);
  dt = 0.99 * courant_length(len_x_global, len_y_global, len_z_global, nx, ny, nz) / c;
  n0 = 1.0;
}
