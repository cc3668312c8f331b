/* bt_oracle_render.c -- CPU oracle, Path B (render).  Filled in below. */
#include "bt_oracle.h"
