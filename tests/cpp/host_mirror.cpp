// Host program over include/clothsim.hpp: runs a reference configuration offline and
// writes one OBJ per cloth per frame (used by the GPU test of the C++ mirror).
//   host_mirror <config.json> <root dir> <out dir> [max frames]
#include <cstdio>
#include <cstdlib>

#include "clothsim.hpp"

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s config root outdir [frames]\n", argv[0]);
        return 2;
    }
    try {
        clothsim::Simulator sim(argv[1], argv[2]);
        const int frames = argc > 4 ? std::atoi(argv[4]) : 2;
        sim.saveFrame(argv[3]);
        while (sim.frames() < frames) {
            const int f = sim.frames();
            cs_stats st = sim.simulateStep();
            if (sim.frames() != f) {
                sim.saveFrame(argv[3]);
                std::printf("frame %d steps %d pcg %d impacts %d\n", sim.frames(), sim.steps(), st.pcg_iters, st.n_impacts);
            }
        }
    } catch (const clothsim::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
