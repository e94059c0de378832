#include "rf_internal.hpp"
