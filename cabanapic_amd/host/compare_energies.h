// compare_energies.h -- the energy-history comparison used by the regression
// deck, with the semantics of tests/energy_comparison/compare_energies.h:
// files of whitespace-separated lines "step time E B ..."; a bit mask picks
// columns (bit k = column k+1); in Sum mode the picked columns of a line are
// added before comparing; a line passes when the relative error
// |A-B|/min(A,B) is below the tolerance, or, when min(A,B) < 10 eps, when
// the absolute error is below 2 eps.  The first skip lines are ignored and
// at most count lines are compared.
#pragma once

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

namespace test_utils {

enum FIELD_ENUM { Individual = 0, Sum };

inline bool within_tolerance(double a, double b, double tol, double* err) {
  const double eps = std::numeric_limits<double>::epsilon();
  const double lo = std::min(a, b);
  if (std::abs(lo) < 10 * eps) {
    *err = std::abs(a - b);
    return *err < 2 * eps;
  }
  *err = std::abs(a - b) / lo;
  return *err < tol;
}

// Returns true when every compared line of file_b matches file_a.  The
// parameters are those of the reference's compare_energies
// (tests/energy_comparison/compare_energies.h:135-145); when write_err_output
// is nonzero the per-line errors (in percent) are written to err_file.
inline bool compare_energies(const std::string& file_a, const std::string& file_b, double tol,
                             unsigned short mask = 0xffffu, FIELD_ENUM mode = Individual,
                             int write_err_output = 0, const std::string& err_file = "err.out",
                             int skip = 0, int count = -1) {
  std::ifstream fa(file_a), fb(file_b);
  if (!fa.is_open()) {
    std::cerr << "Unable to open file " << file_a << std::endl;
    return false;
  }
  if (!fb.is_open()) {
    std::cerr << "Unable to open file " << file_b << std::endl;
    return false;
  }
  std::string la, lb;
  for (int i = 0; i < skip; ++i) {
    std::getline(fa, la);
    std::getline(fb, lb);
  }
  bool ok = true;
  double worst = 0, worst_a = 0, worst_b = 0;
  int worst_line = -1, line = skip, compared = 0;
  std::vector<double> errs;
  while (std::getline(fa, la)) {
    if (!std::getline(fb, lb)) {
      std::cerr << "Second file ends at line " << line + 1 << std::endl;
      ok = false;
      break;
    }
    std::istringstream sa(la), sb(lb);
    double va, vb, suma = 0, sumb = 0;
    bool any = false;
    for (int col = 0; sa >> va; ++col) {
      if (!(sb >> vb)) {
        ok = false;
        break;
      }
      if (col >= 16 || !((mask >> col) & 1u)) continue;
      if (mode == Sum) {
        suma += va;
        sumb += vb;
        any = true;
      } else {
        double e;
        if (!within_tolerance(va, vb, tol, &e)) ok = false;
        errs.push_back(e);
        if (e > worst) worst = e, worst_a = va, worst_b = vb, worst_line = line;
      }
    }
    if (mode == Sum && any) {
      double e;
      if (!within_tolerance(suma, sumb, tol, &e)) ok = false;
      errs.push_back(e);
      if (e > worst) worst = e, worst_a = suma, worst_b = sumb, worst_line = line;
    }
    ++line;
    if (count > 0 && ++compared >= count) break;
  }
  std::cout << "Max found err was " << worst * 100 << "% (" << worst_a << " vs " << worst_b
            << ") on line " << worst_line << " (Threshold: " << tol * 100 << "%)" << std::endl;
  if (write_err_output) {
    std::ofstream out(err_file);
    for (size_t i = 0; i < errs.size(); ++i) out << i + 1 << " " << errs[i] * 100.0 << "\n";
  }
  return ok;
}

}  // namespace test_utils
