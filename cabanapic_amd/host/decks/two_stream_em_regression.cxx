// The reference's energy regression (tests/energy_comparison/2stream-em.cxx):
// an electromagnetic two-stream / filamentation run, 1 x 32 x 1 cells,
// 100 particles per cell from the default initializer, 6000 steps.  After the
// run the energy history is checked against the reference's gold file of the
// same precision with the reference's rule: E energy, then B energy, lines
// 3582-4881, relative error below 10 %.  Exit code 1 on failure.
#include <cmath>
#include <cstdlib>

#include "../compare_energies.h"
#include "../deck.h"

#ifndef GOLD_ENERGY_FILE
#error "GOLD_ENERGY_FILE must name the gold energy history of this precision"
#endif

class Gold_Check : public Run_Finalizer {
 public:
  void finalize() override {
    const char* env = std::getenv("CPIC_GOLD_ENERGY_FILE");
    const std::string gold = env ? env : GOLD_ENERGY_FILE;
    const double margin = 0.10;
    const int skip = 3581, lines = 1300;
    const bool e_ok = test_utils::compare_energies(gold, "energies.txt", margin, 0b100u,
                                                   test_utils::Sum, 1, "e.out", skip, lines);
    std::cout << "E Test Pass: " << e_ok << std::endl;
    const bool b_ok = test_utils::compare_energies(gold, "energies.txt", margin, 0b1000u,
                                                   test_utils::Sum, 1, "b.out", skip, lines);
    std::cout << "B Test Pass: " << b_ok << std::endl;
    if (!e_ok || !b_ok) std::exit(1);
  }
};

Input_Deck::Input_Deck() {
  run_finalizer = new Gold_Check();
  nx = 1;
  ny = 32;
  nz = 1;
  num_steps = 6000;
  nppc = 100;
  v0 = 0.0866025403784439;
  const real_ gam = 1.0 / std::sqrt(1.0 - v0 * v0);
  len_x_global = 1.0;
  len_y_global = 0.628318530717959 * (gam * std::sqrt((double)gam));
  len_z_global = 1.0;
  dt = 0.99 * courant_length(len_x_global, len_y_global, len_z_global, nx, ny, nz) / c;
  n0 = 2.0;  // two beams, each with plasma frequency 1
}
