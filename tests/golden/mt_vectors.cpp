// Writes the libstdc++ <random> draws the reference tests make (mt19937 with
// seeds 3 and 10, uniform_real_distribution<float> and <double>) as raw
// little-endian arrays, for the oracle's restatement of those generators.
//   g++ -O2 -o mt_vectors mt_vectors.cpp && ./mt_vectors out_dir
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

template <class T>
static void dump(const std::string& path, const std::vector<T>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    std::string dir = argc > 1 ? argv[1] : ".";
    {   // raw words, seeds 3 and 10
        for (unsigned seed : {3u, 10u}) {
            std::mt19937 g(seed);
            std::vector<uint32_t> v(2000);
            for (auto& x : v) x = g();
            dump(dir + "/mt_raw_" + std::to_string(seed) + ".u32", v);
        }
    }
    {   // tests/test_closest_neighbors.cu:21-28 and tests/test_frnn.cu:21-28:
        // pos_r(-1000, 1000) float, 2 draws per point, 1000 points
        std::mt19937 g(3);
        std::uniform_real_distribution<float> pos_r(-1000.0f, 1000.0f);
        std::vector<float> v(2000);
        for (auto& x : v) x = pos_r(g);
        dump(dir + "/uniform_f32_seed3.f32", v);
    }
    {   // tests/grid2d_frnn.cu:211-218: dista(-50, 50) double, seed 10, 10000 points
        std::mt19937 g(10);
        std::uniform_real_distribution<> d(-50.0, 50.0);
        std::vector<double> v(20000);
        for (auto& x : v) x = d(g);
        dump(dir + "/uniform_f64_seed10.f64", v);
    }
    {   // the standard's known answer: the 10000th draw of a default-seeded mt19937
        std::mt19937 g;
        uint32_t x = 0;
        for (int i = 0; i < 10000; i++) x = g();
        std::printf("%u\n", x);
    }
    return 0;
}
