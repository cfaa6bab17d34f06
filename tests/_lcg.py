"""The seeded LCG the reference's unit tests draw their inputs from
(tests/test_ao_cholesky.cpp:23, tests/test_jkonly_sigma_hessian_ri.cpp:23-29):
s <- s * 6364136223846793005 + 1442695040888963407 (mod 2^64); value = (s >> 11) / 2^53 - 0.5."""

MASK = (1 << 64) - 1


class Lcg:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & MASK
        return (self.s >> 11) / 9007199254740992.0 - 0.5
