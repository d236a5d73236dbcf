// ParamsCarrier.hpp - key/value parameter store with the interface of the
// reference's ParamsCarrier (Input/include/ParamsCarrier.hpp:22-300), for
// standalone builds of the GPU model adapters.  Inside the reference tree
// the reference's own header is used instead (it comes first on the include
// path); this file only restates its observable behaviour:
//   * one process-wide instance (instance());
//   * putInt / putFloat / putDouble do not replace an existing key
//     (std::map::insert semantics); putString replaces an existing string;
//   * getX(key, default) returns the default when the key is missing or
//     holds another type;
//   * eraseAllItems(except) removes every key not in `except`.
#ifndef GS_COMPAT_PARAMS_CARRIER_HPP
#define GS_COMPAT_PARAMS_CARRIER_HPP

#include <map>
#include <set>
#include <string>

class ParamsCarrier {
  public:
    static ParamsCarrier *instance() {
        static ParamsCarrier *inst = new ParamsCarrier();
        return inst;
    }

    void putString(std::string key, std::string value) {
        auto it = m_.find(key);
        if (it != m_.end() && it->second.type == Type::String) {
            it->second.s = value;
            return;
        }
        Value v;
        v.type = Type::String;
        v.s = value;
        m_.insert({key, v});
    }
    void putFloat(std::string key, float value) {
        Value v;
        v.type = Type::Float;
        v.f = value;
        m_.insert({key, v});
    }
    void putInt(std::string key, int value) {
        Value v;
        v.type = Type::Int;
        v.i = value;
        m_.insert({key, v});
    }
    void putDouble(std::string key, double value) {
        Value v;
        v.type = Type::Double;
        v.d = value;
        m_.insert({key, v});
    }
    int getInt(std::string key, int defaultValue) {
        auto it = m_.find(key);
        return (it != m_.end() && it->second.type == Type::Int) ? it->second.i : defaultValue;
    }
    float getFloat(std::string key, float defaultValue) {
        auto it = m_.find(key);
        return (it != m_.end() && it->second.type == Type::Float) ? it->second.f : defaultValue;
    }
    double getDouble(std::string key, double defaultValue) {
        auto it = m_.find(key);
        return (it != m_.end() && it->second.type == Type::Double) ? it->second.d : defaultValue;
    }
    std::string getString(std::string key, std::string defaultValue) {
        auto it = m_.find(key);
        return (it != m_.end() && it->second.type == Type::String) ? it->second.s : defaultValue;
    }
    void eraseAllItems(std::set<std::string> except) {
        for (auto it = m_.begin(); it != m_.end();) {
            if (except.count(it->first)) ++it;
            else it = m_.erase(it);
        }
    }

  private:
    enum class Type { String, Int, Float, Double };
    struct Value {
        Type type = Type::Int;
        std::string s;
        int i = 0;
        float f = 0.f;
        double d = 0.0;
    };
    ParamsCarrier() = default;
    std::map<std::string, Value> m_;
};

#endif
