// Host-side initial-condition generators: see include/bh_ic.h.
#pragma once
#include "../../../include/bh_ic.h"
