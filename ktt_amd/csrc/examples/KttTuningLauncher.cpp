// Runs a KTT JSON tuning script on the HIP backend: the counterpart of the reference's
// KttTuningLauncher (TuningLoader/TuningLauncher.cpp) for scripts whose Language is "HIP".
//
// Usage: KttTuningLauncher <script.json> [configuration limit] [output file without extension]
#include <cstdio>
#include <string>

#include <ktt_hip.h>

int main(int argc, char** argv)
{
    if (argc < 2)
    {
        std::fprintf(stderr, "usage: %s <script.json> [configuration limit] [output file]\n", argv[0]);
        return 2;
    }
    const long long limit = argc > 2 ? std::stoll(argv[2]) : 0;
    const char* output = argc > 3 ? argv[3] : nullptr;
    char best[2048] = {};
    long long space = 0;
    long long tested = 0;
    long long valid = 0;
    double bestMs = 0.0;
    if (ktt_run_tuning_script(argv[1], output, limit, best, sizeof(best), &space, &tested, &valid, &bestMs) != 0)
    {
        std::fprintf(stderr, "tuning failed: %s\n", ktt_hip_last_error());
        return 1;
    }
    std::printf("space %lld, tested %lld, valid %lld; best %.4f ms: %s\n", space, tested, valid, bestMs, best);
    return valid > 0 ? 0 : 1;
}
