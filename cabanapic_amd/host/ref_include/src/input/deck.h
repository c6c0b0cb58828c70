// The include path of deck sources written for the reference
// (#include "src/input/deck.h"): the deck interface plus the Kokkos and
// Cabana names those sources call on it.
#pragma once
#include "../../../deck.h"
#include "../../../deck_spellings.h"
