# snippet form: evaluated once per cell midpoint (x, y, z in scope)
if x < X:
    rho = 1.0
    ux = 0.0
    p = 1.0
else:
    rho = 0.125
    ux = 0.0
    p = 0.1
