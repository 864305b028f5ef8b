"""Layout of the learned preconditioner's flat weight vector (the order of
cs_set_learned_weights and of the packaged weight file) and its initialisation.

Matrices are [in][out] row-major; every bias follows its matrix.  The two output
layers start at zero, so an untrained network is exactly block Jacobi."""
import numpy as np

H = 64       # hidden features
NLAYER = 3   # message-passing layers
IN = 16      # padded input / output width


def _params():
    p = [("Wn", (IN, H)), ("bn", (H,)), ("We", (IN, H)), ("be", (H,))]
    for l in range(NLAYER):
        p += [("Wa%d" % l, (H, H)), ("Wb%d" % l, (H, H)), ("Wc%d" % l, (H, H)), ("bc%d" % l, (H,)),
              ("W2%d" % l, (H, H)), ("b2%d" % l, (H,)), ("U1%d" % l, (2 * H, H)), ("bu1%d" % l, (H,)),
              ("U2%d" % l, (H, H)), ("bu2%d" % l, (H,))]
    p += [("Dn1", (H, H)), ("bdn1", (H,)), ("Dn2", (H, IN)), ("bdn2", (IN,)),
          ("De1", (H, H)), ("bde1", (H,)), ("De2", (H, IN)), ("bde2", (IN,))]
    return p


PARAMS = _params()
COUNT = sum(int(np.prod(s)) for _, s in PARAMS)
ZERO_INIT = ("Dn2", "bdn2", "De2", "bde2")


def init_weights(seed=0):
    """Uniform(+-1/sqrt(fan_in)) matrices and biases, zero output layers (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    fan_in = 1
    for name, shape in PARAMS:
        if len(shape) == 2:
            fan_in = shape[0]
        if name in ZERO_INIT:
            parts.append(np.zeros(int(np.prod(shape))))
        else:
            b = 1.0 / np.sqrt(fan_in)
            parts.append(rng.uniform(-b, b, int(np.prod(shape))))
    return np.concatenate(parts).astype(np.float32)
