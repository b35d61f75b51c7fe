/*
 * SPDX-License-Identifier: GPL-3.0-or-later
 *
 * This file restates, for testing, algorithms of the risltc renderer
 * (Copyright (C) 2021, Christoph Peters, Karlsruhe Institute of Technology;
 * Copyright (C) 2023, Ishaan Shah, International Institute of Information
 * Technology, Hyderabad), which is licensed under the GNU General Public
 * License version 3 or later. As a derivative of that code, this file is
 * distributed under the same licence: see oracle/COPYING.
 */
/* Exports the static inline helpers of the reference's src/math_utilities.h
 * (compiled from the reference tree by oracle/Makefile into oracle/_ref) so
 * that tests can compare the oracle's restatements against them. Test
 * infrastructure only. */
#include "math_utilities.h"

void ref_matrix_inverse(float inverse[4][4], const float matrix[4][4]) { matrix_inverse(inverse, matrix); }
uint32_t ref_wang_random_number(uint32_t seed) { return wang_random_number(seed); }
float ref_half_to_float(uint16_t half) { return half_to_float(half); }
