"""The backprojector's short correctly rounded square root, reciprocal and
quotient (ddfan_index.h) give the bits of the compiler's correctly rounded sqrtf
and '/' over their whole operand range (tests/native/cr_math_test.hip: every
float in [2^-20, 2^40] for the square root, its reciprocal and the reciprocal
of the float itself, 2e11 quotients over
three full divisor binades and 2^26 random pairs), and the wave form of the
Cephes arcsine gives asinf_cephes's bits on every float bit pattern, and the
fetch position's clamp as v_med3_f32 equals the fminf(fmaxf()) clamp on every
float."""
import subprocess

import pytest

import native_build

pytestmark = pytest.mark.gpu


def test_short_correctly_rounded_sequences_match():
    exe = native_build.cr_math_test()
    print("\n[cr-math] running", exe, flush=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout
    print(out, flush=True)
    assert "cr math ok" in out
