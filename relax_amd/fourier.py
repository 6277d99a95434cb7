"""fourier.f90 on the GPU: normalised in-place 3-D and 2-D real transforms."""
from ._lib import check, f32, load, ptr

FFT_FORWARD, FFT_INVERSE = -1, 1  # fourier.f90:36


def fft3(data, sx1, sx2, sx3, dx1, dx2, dx3, direction, align_data=1):
    """fft3, fourier.f90:177-201 / :330-355 (in place on a (sx1+2)*sx2*sx3 buffer)."""
    check(load().relax_fft3(ptr(f32(data, "data")), sx1, sx2, sx3, dx1, dx2, dx3, direction, align_data), "fft3")


def fft2(data, sx1, sx2, dx1, dx2, direction, align_data=1):
    """fft2, fourier.f90:372-396 / :510-535."""
    check(load().relax_fft2(ptr(f32(data, "data")), sx1, sx2, dx1, dx2, direction, align_data), "fft2")
