// tables.cpp -- host tabulation of quadrature rules and basis functions
// (the constexpr construction in qtable.h, evaluated at run time).
#include "common.h"

int make_qtable(int fe, int rule, QTable *t) { return qt_make(fe, rule, t); }
