"""Parameter layout of the learned-preconditioner U-Net (the layout fx_unet uses)."""

CH = (16, 32, 48, 64)


def layers(cin: int = 2):
    """(cin, cout) of the 15 convolutions in order: encoder 0..3, decoder 2..0, output."""
    out = []
    c = cin
    for l in range(4):
        out += [(c, CH[l]), (CH[l], CH[l])]
        c = CH[l]
    for l in (2, 1, 0):
        out += [(c + CH[l], CH[l]), (CH[l], CH[l])]
        c = CH[l]
    out.append((c, 1))
    return out


def layout(cin: int = 2):
    ls = layers(cin)
    woff, boff = [], []
    o = 0
    for a, b in ls:
        woff.append(o)
        o += 9 * a * b
    for a, b in ls:
        boff.append(o)
        o += b
    return woff, boff, [a for a, _ in ls], [b for _, b in ls]
