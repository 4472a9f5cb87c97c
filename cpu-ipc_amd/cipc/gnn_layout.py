"""Parameter layout of the learned block preconditioner (flat float32 vector, in this
order), shared with the device library (learned.hip): node and edge encoders, three
message-passing layers, node decoder, edge decoder.  Hidden width 64, 16 input features."""
H = 64
F_IN = 16
LAYERS = 3


def param_shapes():
    s = [("Wn", (F_IN, H)), ("bn", (H,)), ("We", (F_IN, H)), ("be", (H,))]
    for l in range(LAYERS):
        s += [(f"Wm{l}", (H, H)), (f"Wu{l}", (2 * H, H)), (f"bu{l}", (H,))]
    s += [("Wd", (H, 16)), ("bd", (16,)), ("Wo", (2 * H, 16)), ("bo", (16,))]
    return s


def n_params():
    n = 0
    for _, shape in param_shapes():
        k = 1
        for d in shape:
            k *= d
        n += k
    return n
