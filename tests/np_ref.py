"""Independent numpy formulation of the multiband blend, used only to pin the
C oracle on small inputs.

It is written from the mathematics rather than from the reference's loops:
pyrDown is a 5x5 [1 4 6 4 1]^2 convolution of the border-padded image sampled
at even positions; pyrUp is zero insertion of the border-padded source
followed by the same 5x5 convolution; distances are brute-force L1 minima.
Sources of truth: Blend/Pyramid.cpp (borders, casts), Blend/MultiBandBlend.cpp
(the blend steps), Blend/Mask.cpp (seams)."""
import numpy as np

K = np.array([1, 4, 6, 4, 1], np.int64)


def pad_hw(img, t, b, l, r):
    """Pad rows with REFLECT_101 (numpy 'reflect') and columns with WRAP."""
    img = np.asarray(img, np.int64)
    pw_rows = ((t, b),) + ((0, 0),) * (img.ndim - 1)
    img = np.pad(img, pw_rows, mode="reflect")
    pw_cols = ((0, 0), (l, r)) + ((0, 0),) * (img.ndim - 2)
    return np.pad(img, pw_cols, mode="wrap")


def conv5(img):
    """Valid 5x5 separable convolution with K x K (int64)."""
    h = sum(K[i] * img[:, i:img.shape[1] - 4 + i] for i in range(5))
    return sum(K[i] * h[i:h.shape[0] - 4 + i] for i in range(5))


def pyr_down_sum(img):
    h, w = img.shape[:2]
    p = pad_hw(img, 2, 2, 2, 2)
    full = conv5(p)            # full[y, x] = centred 5x5 sum at (y, x)
    return full[0:h:2, 0:w:2]


def pyr_down_s16(img):
    return ((pyr_down_sum(img) + 128) >> 8).astype(np.int16)


def pyr_up(img, dshape):
    h, w = img.shape[:2]
    p = pad_hw(img, 1, 1, 1, 1)                   # (h+2, w+2)
    z = np.zeros((2 * (h + 2), 2 * (w + 2)) + img.shape[2:], np.int64)
    z[0::2, 0::2] = p
    full = conv5(z)                                # full[Y, X] centred at z[Y+2, X+2]
    dh, dw = dshape
    return (full[:dh, :dw] + 32) >> 6


def l1_distance(mask):
    """cv::distanceTransform(L1, 3x3) semantics: exact L1 distance to the
    nearest zero pixel, capped by 8192 + distance to the outside frame."""
    h, w = mask.shape
    ys, xs = np.nonzero(mask == 0)
    yy, xx = np.mgrid[0:h, 0:w]
    frame = np.minimum(np.minimum(yy + 1, xx + 1), np.minimum(h - yy, w - xx)) + 8192
    if len(ys) == 0:
        return frame.astype(np.float32)
    d = np.abs(yy[..., None] - ys) + np.abs(xx[..., None] - xs)
    return np.minimum(d.min(axis=-1), frame).astype(np.float32)


def unique_masks(masks):
    n = len(masks)
    out = [m.copy() for m in masks]
    acc = out[0].copy()
    for i in range(1, n):
        d = l1_distance(acc)
        cur = l1_distance(out[i])
        out[i] = np.where(cur > d, 255, 0).astype(np.uint8)
        for j in range(i):
            out[j][out[i] != 0] = 0
        acc |= out[i]
    return np.stack(out)


def num_levels(w, h, max_levels, min_len=2):
    m = min(int(np.log2(w / min_len)), int(np.log2(h / min_len)))
    return min(max_levels, m)


def tdiv(a, b):
    """C truncating division on int64 arrays."""
    q = np.abs(a) // np.abs(b)
    return np.where((a < 0) != (b < 0), -q, q)


def multiband(images, masks, max_levels, min_len=2, weight_masks=None):
    """images: (n, H, W, 3) int16; masks (n, H, W) u8.  Returns (H, W, 3) u8.
    weight_masks: the custom-mask blend (MultiBandBlend.cpp:1033-1102): weights
    from these masks instead of the seams, always divided by their sums, and
    the output zeroed outside their union."""
    n, H, W = masks.shape
    L = num_levels(W, H, max_levels, min_len)
    uniq = unique_masks(masks) if weight_masks is None else weight_masks
    sizes = [(H, W)]
    for _ in range(L):
        sizes.append(((sizes[-1][0] + 1) // 2, (sizes[-1][1] + 1) // 2))
    weights, alphas = [], []
    for i in range(n):
        wp = [np.where(uniq[i] != 0, 256, 0).astype(np.int16)]
        for _ in range(L):
            wp.append(pyr_down_s16(wp[-1]))
        weights.append(wp)
        ap = [None]
        cur = np.where(masks[i] != 0, 256, 0)
        for _ in range(L):
            a = pyr_down_sum(cur)
            ap.append(a)
            cur = np.where(a != 0, 256, 0)
        alphas.append(ap)
    cover = np.any((masks if weight_masks is None else weight_masks) != 0, axis=0)
    full = bool(cover.all())
    R = [np.zeros(s + (3,), np.int64) for s in sizes]
    for i in range(n):
        G = [images[i].astype(np.int64)]
        for j in range(L):
            s = pyr_down_sum(G[j])
            a = alphas[i][j + 1][..., None]
            v = (s * 255)
            G.append(np.where(a != 0, tdiv(v, np.where(a == 0, 1, a)), 0).astype(np.int16).astype(np.int64))
        Lp = []
        for j in range(L):
            up = pyr_up(G[j + 1], sizes[j]).astype(np.int16).astype(np.int64)
            Lp.append(np.clip(G[j] - up, -32768, 32767))
        Lp.append(G[L])
        for j in range(L + 1):
            R[j] += Lp[j] * weights[i][j].astype(np.int64)[..., None]
    for j in range(L + 1):
        if full and weight_masks is None:
            R[j] = R[j] >> 8
        else:
            wsum = sum(weights[i][j].astype(np.int64) for i in range(n))
            wsum = np.where(wsum == 0, 1, wsum)[..., None]
            R[j] = tdiv(R[j], wsum)
    for j in range(L, 0, -1):
        R[j - 1] = R[j - 1] + pyr_up(R[j], sizes[j - 1])
    out = np.clip(R[0], 0, 255).astype(np.uint8)
    if not full:
        out[~cover] = 0
    return out
