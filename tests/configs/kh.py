# init_global form: whole arrays of shape (nx, ny, nz), x the fastest index.
# Band rho = 2, ux = -0.5 between the perturbed interfaces 0.25 + d0(x) and
# 0.75 + d1(x), d(x) = 0.01 sum_i a_i cos(2 pi ((i+1) x + b_i)) / sum(a);
# parameters a[0:10] / a[20:30] perturb the upper interface, a[10:20] /
# a[30:40] the lower one.
def init_global(rho, ux, uy, p, nx, ny, nz, ax, ay, az, bx, by, bz):
    dx = (bx - ax) / nx
    dy = (by - ay) / ny
    x = ax + dx * arange(nx) + dx / 2.0
    y = ay + dy * arange(ny) + dy / 2.0

    def delta(amp, phase):
        norm = sum(amp)
        if abs(norm) < 1e-10:
            norm = 1.0
        out = zeros_like(x)
        for i in range(len(amp)):
            out = out + amp[i] * cos(2 * pi * ((i + 1) * x + phase[i]))
        return 0.01 * out / norm

    hi = 0.75 + delta(a[0:10], a[20:30])
    lo = 0.25 + delta(a[10:20], a[30:40])
    Y = y[None, :]
    band = (Y > lo[:, None]) & (Y < hi[:, None])
    rho[:, :, 0] = where(band, 2.0, 1.0)
    ux[:, :, 0] = where(band, -0.5, 0.5)
    uy[:, :, 0] = 0.0
    p[:, :, 0] = 2.5
