// Driver for the reference-compiled checker (test infrastructure only).
// Links nothing but the reference's own headers REXI.hpp,
// GaussianApproximation.hpp, ExponentialApproximation.hpp (src/include/rexi/)
// and SPHIdentities.hpp (src/include/sph/), compiled from /root/reference.
//
// usage: ref_rexi_dump rexi <h> <M> <half>     -> "re_alpha im_alpha re_beta_re im_beta_re re_beta_im im_beta_im"
//                                                  per pole
//        ref_rexi_dump ident <L>               -> "n m R S A B C D E G H" for 0<=m<=n<=L (and m<0);
//                                                 H(0, m) is printed as nan (the reference asserts there)
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <rexi/REXI.hpp>
#include <sph/SPHIdentities.hpp>

int main(int argc, char **argv)
{
	if (argc >= 5 && std::strcmp(argv[1], "rexi") == 0) {
		REXI r;
		r.setup(std::atof(argv[2]), std::atoi(argv[3]), 0, std::atoi(argv[4]) != 0);
		for (std::size_t i = 0; i < r.alpha.size(); i++)
			std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", r.alpha[i].real(), r.alpha[i].imag(),
				    r.beta_re[i].real(), r.beta_re[i].imag(), r.beta_im[i].real(), r.beta_im[i].imag());
		return 0;
	}
	if (argc >= 3 && std::strcmp(argv[1], "ident") == 0) {
		int L = std::atoi(argv[2]);
		for (int n = 0; n <= L; n++)
			for (int m = -n; m <= n; m++)
				std::printf("%d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", n, m,
					    SPHIdentities::R(n, m), SPHIdentities::S(n, m), SPHIdentities::A(n, m),
					    SPHIdentities::B(n, m), SPHIdentities::C(n, m),
					    n + 1 >= 0 ? SPHIdentities::D(n, m) : 0.0, SPHIdentities::E(n, m),
					    SPHIdentities::G(n, m),
					    n >= 1 ? SPHIdentities::H(n, m) : NAN);   // H(0, m) asserts on E(-1, m)
		return 0;
	}
	std::fprintf(stderr, "usage: %s rexi <h> <M> <half> | ident <L>\n", argv[0]);
	return 1;
}
