// Single include of the host mirror of DaRe's FV hot-path call surface on MI355X.
#pragma once
#include "Device.h"
#include "Grid.h"
#include "Data.h"
#include "MatrixSystem.h"
#include "IO.h"
