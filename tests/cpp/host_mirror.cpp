// Host program over include/clothsim.hpp: runs a reference configuration offline, saving every
// frame as the reference does (frame<N>_cloth<i>.obj + config.json), or resumes such a run.
//   host_mirror <config.json> <root dir> <out dir> [frames [overlapped]]
//   host_mirror --resume <out dir> <root dir> [frames]
// With "overlapped", frames after the first are written by Simulator::saveOverlapped.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "clothsim.hpp"

int main(int argc, char** argv) {
    const bool resume = argc > 1 && std::strcmp(argv[1], "--resume") == 0;
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s config root outdir [frames] | --resume outdir root [frames]\n", argv[0]);
        return 2;
    }
    try {
        std::unique_ptr<clothsim::Simulator> sim;
        std::string out;
        int frames = 2;
        bool overlapped = false;
        if (resume) {
            out = argv[2];
            sim.reset(new clothsim::Simulator(clothsim::Resume{}, out, argv[3]));
            frames = argc > 4 ? std::atoi(argv[4]) : sim->frames() + 1;
            std::printf("resumed at frame %d step %d\n", sim->frames(), sim->steps());
        } else {
            out = argv[3];
            sim.reset(new clothsim::Simulator(argv[1], argv[2]));
            frames = argc > 4 ? std::atoi(argv[4]) : 2;
            overlapped = argc > 5 && std::strcmp(argv[5], "overlapped") == 0;
            sim->save(out);
        }
        while (sim->frames() < frames) {
            cs_stats st = sim->simulateStep();
            if (sim->steps() % sim->info().frame_steps == 0) {
                if (overlapped) sim->saveOverlapped(out);
                else sim->save(out);
                std::printf("frame %d steps %d pcg %d impacts %d\n", sim->frames(), sim->steps(), st.pcg_iters,
                            st.n_impacts);
            }
        }
        sim->flush();
    } catch (const clothsim::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
