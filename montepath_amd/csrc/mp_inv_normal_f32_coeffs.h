// Generated by gen_inv_normal_f32.py: central branch of mp_inv_normal_f32,
// R(l) = sum c_i l^i, l = log2(1 - x^2), for -ln(1 - x^2) < 7.0.
#pragma once
#define MP_INVN_L_CENTRAL -10.098865509033203f  // log2(1 - x^2) above this: w < 7.0
#define MP_INVN_C0 1.2533138990402222f
#define MP_INVN_C1 -0.22743871808052063f
#define MP_INVN_C2 0.00782822910696268f
#define MP_INVN_C3 0.0010744904866442084f
#define MP_INVN_C4 -5.802952728117816e-05f
#define MP_INVN_C5 -1.6110017895698547e-05f
#define MP_INVN_C6 -6.674715109511453e-07f
#define MP_INVN_C7 5.386894841308276e-08f
#define MP_INVN_C8 5.8498521582350804e-09f
#define MP_INVN_C9 1.5682127818550384e-10f
