// beam.h - Beam is declared with Bunch (bunch.h); this header keeps the
// reference's include name (src/beam.h).
#pragma once

#include "bunch.h"
