// Single include of the host mirror of DaRe's FV hot-path call surface on MI355X.
#pragma once
#include "Vector.h"
#include "Device.h"
#include "Grid.h"
#include "Data.h"
#include "Stencils.h"
#include "MatrixSystem.h"
#include "Operators.h"
#include "IO.h"
