// Magnetized electron column (a 3D version of the reference's
// decks/dioctron_3d.cxx): 192 x 192 x 64 periodic cells of edge 0.1, 32
// electrons per cell (75,497,472) whose weights carry the ring density
// n(r) = 0.25 exp(-((r - 0.3 Lx) / 0.05 Lx)^2) (1 + 1e-3 cos(3 theta)) about
// the z axis through the box centre, guide field B_z = 1 (peak
// w_ce / w_pe = 2), E x B drift in the column's own radial field, which also
// initializes ex, ey in every voxel; seed 3, dt = 0.99 Courant, 1000 steps.
#include "../plasma_init.h"

Input_Deck::Input_Deck() {
  auto* p = new Plasma_Initializer();
  auto* f = new Column_Field_Initializer();
  f->plasma = p;
  particle_initer = p;
  field_initer = f;
  nx = ny = 192;
  nz = 64;
  nppc = 32;
  num_steps = 1000;
  len_x_global = len_y_global = (real_)(192 * 0.1);
  len_z_global = (real_)(64 * 0.1);
  dt = 0.99 * courant_length(len_x_global, len_y_global, len_z_global, nx, ny, nz) / c;
  n0 = 0.25;
  p->seed = 3;
  p->pert_amp = 1e-3;
  p->pert_mode = 3;
  p->column = true;
  p->ring_radius = 0.3 * (double)len_x_global;
  p->ring_width = 0.05 * (double)len_x_global;
  p->column_charge = (double)qsp * 0.25;
  p->guide_bz = 1.0;
}
