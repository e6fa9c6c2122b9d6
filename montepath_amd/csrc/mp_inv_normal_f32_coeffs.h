// Generated by gen_inv_normal_f32.py: central branch of mp_inv_normal_f32,
// R(t) = sum c_i t^i, t = w - 3.5, w = -ln(1 - x^2) < 7.0.
#pragma once
#define MP_INVN_L_CENTRAL -10.098865509033203f  // log2(1 - x^2) above this: w < 7.0
#define MP_INVN_CENTER 3.5f
#define MP_INVN_C0 2.4647340774536133f
#define MP_INVN_C1 0.3328462839126587f
#define MP_INVN_C2 -0.009501068852841854f
#define MP_INVN_C3 -0.0006797267124056816f
#define MP_INVN_C4 0.0002254102728329599f
#define MP_INVN_C5 -2.2656324290437624e-05f
#define MP_INVN_C6 -8.2237318110856e-07f
#define MP_INVN_C7 5.007536287848779e-07f
#define MP_INVN_C8 -2.3963178108488137e-08f
#define MP_INVN_C9 -4.245959139836941e-09f
