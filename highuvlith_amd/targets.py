"""ILT targets and pattern-error counts (crates/highuvlith-core/src/ilt.rs:268-362)."""
import numpy as np


def pattern_error(image, target, threshold):
    """Pixels whose printed state (image > threshold) differs from target > 0.5 (ilt.rs:269-275)."""
    return int(np.count_nonzero((np.asarray(image) > threshold) != (np.asarray(target) > 0.5)))


def pixel_centres(grid):
    half = grid.field_size_nm() / 2.0
    return -half + (np.arange(grid.size) + 0.5) * grid.pixel_nm


def create_target_line_space(cd_nm, pitch_nm, grid):
    """Bright spaces of width pitch - cd, dark lines of width cd (ilt.rs:279-297)."""
    x = pixel_centres(grid)
    x_mod = np.mod(x / pitch_nm + 0.5, 1.0)
    row = (x_mod < (pitch_nm - cd_nm) / pitch_nm).astype(np.float64)
    return np.tile(row, (grid.size, 1))


def create_target_from_fn(grid, supersample, inside):
    """Area-averaged indicator on supersample^2 sub-pixel samples (ilt.rs:300-325).

    ``inside(x, y)`` receives numpy arrays (nm) and returns a boolean array.
    """
    n, p = grid.size, grid.pixel_nm
    half = grid.field_size_nm() / 2.0
    s = max(int(supersample), 1)
    sub = (np.arange(s) + 0.5) / s
    ys = -half + (np.arange(n)[:, None] + sub[None, :]) * p  # (n, s)
    xs = ys
    hits = np.zeros((n, n))
    for a in range(s):
        y = ys[:, a][:, None]
        for b in range(s):
            x = xs[:, b][None, :]
            hits += inside(np.broadcast_to(x, (n, n)), np.broadcast_to(y, (n, n)))
    return hits / (s * s)


def create_target_contact_array(size_nm, pitch_x_nm, pitch_y_nm, count_x, count_y, shape, grid):
    """Centred count_x x count_y array of square or round holes (ilt.rs:336-362)."""
    centres = [((ix - (count_x - 1) / 2.0) * pitch_x_nm, (iy - (count_y - 1) / 2.0) * pitch_y_nm)
               for iy in range(count_y) for ix in range(count_x)]
    return create_target_holes(grid, centres, size_nm, shape)


def create_target_holes(grid, centres, size_nm, shape="square"):
    """Holes of side / diameter ``size_nm`` at arbitrary centres (4x supersampled)."""
    r = size_nm / 2.0
    sizes = np.broadcast_to(np.asarray(size_nm, dtype=np.float64), (len(centres),))

    def inside(x, y):
        acc = np.zeros(x.shape, dtype=bool)
        for (cx, cy), sz in zip(centres, sizes):
            rr = sz / 2.0
            if shape == "square":
                acc |= (np.abs(x - cx) <= rr) & (np.abs(y - cy) <= rr)
            else:
                acc |= (x - cx) ** 2 + (y - cy) ** 2 <= rr * rr
        return acc

    del r
    return create_target_from_fn(grid, 4, inside)
