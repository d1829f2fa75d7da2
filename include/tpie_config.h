// Build configuration of the TPIE subset (the reference's tpie_config.h
// tunes TPIE block and stream sizes; this subset has nothing to tune).
#ifndef TERRANNI_TPIE_CONFIG_H
#define TERRANNI_TPIE_CONFIG_H
#include "tpie/tpie.h"
#endif
