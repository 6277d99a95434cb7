"""okada/green_space.f90 on the GPU: Okada's (1992) displacement of a
rectangular dislocation in a half-space on the grid (grnfct_okada)."""
from ._lib import check, f32, load, ptr


def grnfct_okada(lam, mu, slip, opening, x, y, z, width, length, strike, dip, rake, dx1, dx2, dx3, u1, u2, u3,
                 align_data=1):
    """grnfct_okada, green_space.f90:119-146: adds the displacement of one
    fault (upper corner x, y, z at the start of the strike, north-east-down;
    strike, dip, rake in degrees) to u1..u3, (sx3, sx2, sx1+2) float32 arrays
    (align_data=1) or (sx3, sx2, sx1) (align_data=0), in place."""
    sx3, sx2, p = u1.shape
    sx1 = p - 2 if align_data else p
    check(load().relax_grnfct_okada(lam, mu, slip, opening, x, y, z, width, length, strike, dip, rake, sx1, sx2, sx3,
                                    dx1, dx2, dx3, ptr(f32(u1, "u1")), ptr(f32(u2, "u2")), ptr(f32(u3, "u3")),
                                    align_data), "grnfct_okada")
