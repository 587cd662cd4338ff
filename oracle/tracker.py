"""CPU restatement of ins::tracker's step-size rule (ins/tracker.h:13-100);
test infrastructure only.

  |f|   = sqrt(sum over force arrays of max|f|^2)            (nrminf2, :42-49)
  k_max = (1/4) sqrt(fscale / (|f| + 1e-3))                  (bound, :62-75)
  e     = floor(log2(k_max / k_min))                         (serialize, :51)
  dt    = k_min 2^min(4, e over the last 16 steps incl. this) (current, :53-60, emax :15)
  e < 0 raises too_small                                     (:93)
The retry loop of :84-91 compares the newest exponent with a minimum that
already contains it and never runs, so it has no counterpart.
"""
import math


class TooSmall(RuntimeError):
    pass


class Tracker:
    def __init__(self, kmin, fscale):
        self.kmin, self.fscale = kmin, fscale
        self.history = []

    def bound(self, forces):
        fmax = math.sqrt(sum(float(abs(f).max()) ** 2 for f in forces if f.size))
        kmax = 0.25 * math.sqrt(self.fscale / (fmax + 1e-3))
        return math.floor(math.log2(kmax / self.kmin))

    def __call__(self, forces):
        self.history = (self.history + [self.bound(forces)])[-16:]
        e = min([4] + self.history)
        if e < 0:
            raise TooSmall()
        return self.kmin * 2 ** e, e
