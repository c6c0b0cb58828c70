// The include path of the reference's regression deck
// (#include "tests/energy_comparison/compare_energies.h").
#pragma once
#include "../../../compare_energies.h"
