// afm_dropin.hpp -- every drop-in header at once.  The headers live at the
// reference's include paths (include/dt, alg, mg, ns, ib), so programs
// written for the reference compile against them with -I include; see
// afm/dropin_core.hpp for what runs where.
#pragma once
#include "dropin_core.hpp"
#include "../dt/geometry.h"
#include "../dt/vector.h"
#include "../dt/point.h"
#include "../dt/matrix.h"
#include "../dt/kdtree.h"
#include "../alg/cg.h"
#include "../alg/smooth.h"
#include "../alg/div.h"
#include "../alg/grad.h"
#include "../mg/laplacian.h"
#include "../mg/solver.h"
#include "../ns/helmholtz.h"
#include "../ns/projection.h"
#include "../ns/solver.h"
#include "../ib/grid.h"
#include "../ib/eulerian_grid.h"
#include "../ib/mac_grid.h"
#include "../ib/lagrangian_grid.h"
