"""MI355X (gfx950) tomographic two-point projection: Limber C_ell (ss, gs, gg)
and bin-averaged Legendre transforms (xi+-, gamma_t, w), as a drop-in for
CosmoLike's real-space 3x2pt path."""
from .interface import Cosmolike, CosmolikeError, Layout, LEGENDRE_FOLDED, LEGENDRE_FULL, PRECISION  # noqa: F401
