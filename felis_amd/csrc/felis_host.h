// Host helpers shared by the deck, loader and run-loop translation units.
#pragma once

#include "felis.h"
#include "felis_deck.h"

int felis_deck_parse(const char *path, felis_deck *D);
void felis_update_k_quadg(const felis_deck *D, felis_lattice *lat, int iteration, double half);
felis_lattice felis_initial_lattice(const felis_deck *D);
int felis_wake_tables(const felis_deck *D, double *wakeF, double *impZ_out);
