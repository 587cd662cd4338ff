// Declarations in the shapes reference programs use (test/grid.cc:28-35,
// test/ib-whole.cc:56-63,197-199), compiled unchanged against include/ibm.hpp.
// The unit values are checked against the reference's scales
// (units.h:16-18: 1000 length units per cm, 10^6 mass units per g, 1000
// time units per s).  Compile-only: nothing here touches a device.
#include "ibm.hpp"

constexpr bool near(double a, double b) { return (a - b) <= 1e-15 * b && (b - a) <= 1e-15 * b; }

int main()
{
    // test/grid.cc:28-35
    constexpr fd::dimension x{10_um, fd::boundary::periodic()};
    constexpr fd::dimension y{10_um, fd::boundary::periodic()};
    constexpr fd::dimension z{10_um, fd::boundary::periodic()};
    constexpr fd::domain domain{x, y, z};
    constexpr fd::mac mac{16};
    constexpr ib::spread spread{mac, domain};
    constexpr ib::interpolate interpolate{mac, domain};
    static_assert(near(x.length(), 1.0));

    // test/ib-whole.cc:197-199: objects rather than constructor calls
    constexpr fd::dimension wx{16_um, fd::boundary::periodic};
    constexpr fd::dimension wy{12_um, fd::boundary::dirichlet};
    constexpr fd::dimension wz{16_um, fd::boundary::periodic};
    constexpr fd::domain walled{wx, wy, wz};
    static_assert(wy.d.solid == 1 && wy.d.lower[0] == 1.0 && wy.d.lower[1] == 0.0);
    static_assert(wx.d.solid == 0);
    constexpr fd::dimension nz{16_um, fd::boundary::neumann()};
    static_assert(nz.d.lower[0] == 0.0 && nz.d.lower[1] == 1.0);
    constexpr fd::dimension rb{16_um, fd::boundary::robin(2.0, 0.5), fd::boundary::dirichlet};
    static_assert(rb.d.lower[0] == 2.0 && rb.d.upper[0] == 1.0);

    // test/ib-whole.cc:56-63 style quantities
    units::time time = 0_s;
    constexpr units::hertz rate = 1000 / 1_s;
    constexpr units::diffusivity nu = 1_cP / (1_g / 1_mL);
    static_assert(1_s == 1000.0 && 1_ms == 1.0 && near(100_ns, 1e-4));
    static_assert(1_cm == 1000.0 && 1_mm == 100.0 && near(1.5_um, 0.15) && near(12_um, 1.2));
    static_assert(1_g == 1e6 && 1_kg == 1e9);
    static_assert(nu > 9.999999999 && nu < 10.000000001);   // 1e-6 m^2/s
    static_assert(rate == 1.0);
    static_assert(1_dyn == 1e3 && 1_erg == 1e6 && 1_P == 1.0 && near(1_N, 1_kg * 1_m / (1_s * 1_s)));
    constexpr units::energy kappa = 2e-19_J;
    static_assert(kappa > 1.99e-6 && kappa < 2.01e-6);
    (void) time;
    (void) spread;
    (void) interpolate;
    (void) walled;
    return 0;
}
