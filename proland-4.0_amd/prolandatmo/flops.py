"""Algorithmic work of the atmo precompute, counted on the reference's own
formulation (the GLSL passes of src/atmo/sources/proland/preprocess/atmo/).

Convention: every FP add/sub/mul/div/min/max/compare/select/floor and every
transcendental (sqrt, exp, pow, sin, cos, tan, atan) is 1 FLOP; per-set
constants (e.g. tan(1.26*1.1), 1/RES_MU, 3/(16 pi)) are folded; a GL LINEAR
fetch counts its coordinate arithmetic (u*N - 0.5, floor, frac per axis) and
its lerps (3 FLOPs per channel each); vec3 ops count 3.  These are the FLOPs a
literal evaluation of the reference shaders performs, independent of how the
kernels here reorganise the work (they stage the separable part of the
filtering once per workgroup and so execute far fewer).
"""

# --- building blocks (common.glsl) -------------------------------------------
FETCH2D_RGB = 2 * 4 + 3 * 3 * 3           # 2 axes x 4 + 3 lerps x 3 ch x 3   = 35
FETCH3D_RGB = 3 * 4 + 7 * 3 * 3           # 3 axes x 4 + 7 lerps x 3 ch x 3   = 75

TEXTURE4D = {                              # common.glsl:149-173
    "rho = sqrt(r*r - Rg*Rg)": 3,
    "rmu = r*mu": 1,
    "delta = rmu*rmu - r*r + Rg*Rg": 4,
    "cst = rmu < 0 && delta > 0 ? ... (vec4 select)": 7,
    "uR": 3,
    "uMu": 8,
    "uMuS (atan)": 7,
    "lerp, uNu = floor, frac": 5,
    "two x coordinates": 5,
    "two trilinear RGB fetches": 2 * FETCH3D_RGB,
    "nu blend (vec4)": 13,
}
T_LOOKUP = {                               # common.glsl:114-124, 223-226
    "uR = sqrt((r-Rg)/(Rt-Rg))": 3,
    "uMu = atan((mu+0.15)/1.15*tan(1.5))/1.5": 5,
    "bilinear RGB fetch": FETCH2D_RGB,
}
T_SEGMENT = {                              # common.glsl:272-282
    "r1 = sqrt(r*r + d*d + 2*r*mu*d)": 8,
    "mu1 = (r*mu + d)/r1": 3,
    "mu > 0 test": 1,
    "two T lookups": 2 * sum(T_LOOKUP.values()),
    "negation": 1,
    "ratio and min(., 1)": 6,
}
IRRADIANCE_LOOKUP = {                      # common.glsl:138-142, 284-287
    "uR": 2, "uMuS": 2, "bilinear RGB fetch": FETCH2D_RGB,
}
PHASE_R = 3                                # common.glsl:290-292
PHASE_M = 8                                # common.glsl:295-297 (pow counted once)
GET_MU_MUS_NU = 35                         # common.glsl:175-201
LIMIT = 14                                 # common.glsl:209-219

TEX4D = sum(TEXTURE4D.values())            # 206
TLOOK = sum(T_LOOKUP.values())             # 43
TSEG = sum(T_SEGMENT.values())             # 105
ELOOK = sum(IRRADIANCE_LOOKUP.values())    # 39

# --- inscatterS.glsl:111-189, per direction (theta, phi) ----------------------
INSCATTERS_DIR_COMMON = {
    "phi": 2, "dw": 2, "w": 6, "nu1 = dot(s,w)": 5, "nu2 = dot(v,w)": 5,
    "pr2": PHASE_R, "pm2": PHASE_M, "gnormal": 9, "dot(gnormal, s)": 5,
    "girradiance": ELOOK, "raymie1 = greflectance*girradiance*gtransp": 6,
    "raymie += raymie1 * (betaR e^-h/HR pr2 + betaMSca e^-h/HM pm2) * dw": 30,
}
INSCATTERS_DIR_LATER = {"texture4D(dSR)": TEX4D, "raymie1 +=": 3}
INSCATTERS_DIR_FIRST = {"pr1": PHASE_R, "pm1": PHASE_M, "texture4D(dSR), texture4D(dSM)": 2 * TEX4D,
                        "ray1*pr1 + mie1*pm1": 9, "raymie1 +=": 3}
INSCATTERS_THETA = 4 + (8 + 3 + TSEG) // 2  # theta, cos, test; ground branch on ~half the rows
INSCATTERS_TEXEL = GET_MU_MUS_NU + 32       # getMuMuSNu + clamps, v, s

# --- the other passes, per texel ---------------------------------------------
TRANSMITTANCE_SAMPLE = 15                  # transmittance.glsl:83-89
INSCATTER1_SAMPLE = 13 + 1 + 6 + TSEG + TLOOK + 3 + 12 + 12   # inscatter1.glsl:104-135
INSCATTERN_SAMPLE = 13 + TEX4D + TSEG + 3 + 12 + 1            # inscatterN.glsl:105-123
IRRADIANCEN_DIR_LATER = 2 + 2 + 2 + 6 + 5 + TEX4D + 9        # irradianceN.glsl:92-111
IRRADIANCEN_DIR_FIRST = IRRADIANCEN_DIR_LATER + TEX4D + PHASE_R + PHASE_M + 9
IRRADIANCE1_TEXEL = 8 + TLOOK + 1 + 3
COPY_INSCATTERN_TEXEL = GET_MU_MUS_NU + PHASE_R + 3 + 4


def inscatterS_dir(first):
    return sum(INSCATTERS_DIR_COMMON.values()) + sum(
        (INSCATTERS_DIR_FIRST if first else INSCATTERS_DIR_LATER).values())


def inscatterS_texel(ns, first):
    """FLOPs of one dJ texel with an ns x 2ns direction grid."""
    return 2 * ns * ns * inscatterS_dir(first) + ns * INSCATTERS_THETA + INSCATTERS_TEXEL


def per_pass(p, o):
    """{pass: FLOPs of all its launches} for one parameter set p with options o."""
    n4 = p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU
    nT = p.TRANSMITTANCE_W * p.TRANSMITTANCE_H
    nE = p.SKY_W * p.SKY_H
    ns, ni, orders = o.spherical_samples, o.irradiance_samples, o.orders
    s1 = o.inscatter_samples + 1
    w = {
        "transmittance": nT * (2 * (LIMIT + 1 + TRANSMITTANCE_SAMPLE * o.transmittance_samples + 5) + 12),
        "irradiance1": nE * IRRADIANCE1_TEXEL,
        "inscatter1": n4 * (GET_MU_MUS_NU + LIMIT + 1 + s1 * INSCATTER1_SAMPLE + 6),
        "inscatterS": 0, "irradianceN": 0, "inscatterN": 0,
    }
    ndirE = 2 * ni * (ni // 2)
    for n in range(2, orders + 1):
        first = n == 2
        w["inscatterS"] += n4 * inscatterS_texel(ns, first)
        w["irradianceN"] += nE * ndirE * (IRRADIANCEN_DIR_FIRST if first else IRRADIANCEN_DIR_LATER)
        w["inscatterN"] += n4 * (GET_MU_MUS_NU + LIMIT + 1 + s1 * INSCATTERN_SAMPLE +
                                 COPY_INSCATTERN_TEXEL)
    return w


def total(p, o):
    return sum(per_pass(p, o).values())
