// Prefix: a sorted list of earlier pattern vertices whose neighbour lists are
// intersected (reference include/prefix.h:3-25).  has_child and
// only_need_size survive init(), as in the reference (src/prefix.cpp:18-23):
// Schedule_IEP reuses its prefix table across candidate orders and the final
// flags depend on that history.
#pragma once
#include <vector>

class Prefix {
public:
    Prefix() : has_child(false), only_need_size(false) {}
    void init(int input_size, const int* input_data) { data.assign(input_data, input_data + input_size); }
    bool equal(int input_size, const int* input_data) const {
        if ((int)data.size() != input_size) return false;
        for (int i = 0; i < input_size; ++i)
            if (data[i] != input_data[i]) return false;
        return true;
    }
    bool operator==(const Prefix& o) const { return data == o.data; }
    bool operator!=(const Prefix& o) const { return data != o.data; }
    inline int get_size() const { return (int)data.size(); }
    inline const int* get_data_ptr() const { return data.data(); }
    inline int get_data(int i) const { return data[i]; }
    void set_has_child(bool f) { has_child = f; }
    bool get_has_child() const { return has_child; }
    void set_only_need_size(bool f) { only_need_size = f; }
    bool get_only_need_size() const { return only_need_size; }

private:
    std::vector<int> data;
    bool has_child;
    bool only_need_size;
};
